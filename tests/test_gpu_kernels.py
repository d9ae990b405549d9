"""Kernel-level parity through the C ABI (vj_op_*): the reference's CUDA kernel tests
(tests/test_cuda_kernels.cu:74-359) restated with their vectors, plus SpMV against the
oracle on the seeded 2,000 x 5,000 LP."""
import numpy as np
import pytest

import lp_cases
import oracle as O

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def dev(a, dtype=None):
    a = np.ascontiguousarray(a)
    t = torch.from_numpy(a)
    if dtype is not None:
        t = t.to(dtype)
    return t.cuda()


def test_spmv_csr_kat():
    # test_cuda_kernels.cu:74-107 -> {-4.5, 0, 14, 11}, empty row exactly 0
    import vajra_pdhg as vp
    rp = dev(np.array([0, 2, 2, 5, 6], np.int64))
    ci = dev(np.array([0, 2, 0, 3, 4, 1], np.int32))
    v = dev(np.array([1.5, -2.0, 3.0, 4.0, -1.0, 5.5]))
    x = dev(np.array([1.0, 2.0, 3.0, 4.0, 5.0]))
    for lanes, exact in [(2, False), (4, False), (8, False), (32, False), (64, False), (1, True)]:
        y = torch.full((4,), 7.0, dtype=torch.float64, device="cuda")
        vp.op_spmv(rp, ci, v, x, y, lanes=lanes, exact_order=exact)
        torch.cuda.synchronize()
        np.testing.assert_allclose(y.cpu().numpy(), [-4.5, 0.0, 14.0, 11.0], rtol=1e-12, atol=0)
        assert y.cpu().numpy()[1] == 0.0


def test_spmv_transpose_kat():
    # test_cuda_kernels.cu:110-140: CSC of a 4x3 matrix as CSR of A^T -> {7, 0, 10}
    import vajra_pdhg as vp
    cp = dev(np.array([0, 2, 2, 4], np.int64))
    ri = dev(np.array([0, 2, 1, 3], np.int32))
    v = dev(np.array([1.0, 2.0, -1.0, 3.0]))
    y = dev(np.array([1.0, 2.0, 3.0, 4.0]))
    out = torch.zeros(3, dtype=torch.float64, device="cuda")
    vp.op_spmv(cp, ri, v, y, out, lanes=4)
    torch.cuda.synchronize()
    np.testing.assert_allclose(out.cpu().numpy(), [7.0, 0.0, 10.0], rtol=1e-12, atol=0)


def test_dual_update_kat():
    # test_cuda_kernels.cu:143-173: sigma = 0.5 -> {1, 0, -1, 0.5}
    import vajra_pdhg as vp
    w = dev(np.array([10.0, 5.0, 2.0, 7.0]))
    lr = dev(np.array([8.0, -1e20, 4.0, 3.0]))
    ur = dev(np.array([8.0, 6.0, 1e20, 6.0]))
    y = dev(np.zeros(4))
    vp.op_dual_update(0.5, w, lr, ur, y)
    torch.cuda.synchronize()
    np.testing.assert_allclose(y.cpu().numpy(), [1.0, 0.0, -1.0, 0.5], rtol=1e-12, atol=0)
    # test_pdhg.cpp:479-497 projections (==, <=, >=, ranged)
    w = dev(np.array([10.0, 8.0, 2.0, 3.0]))
    lr = dev(np.array([5.0, -1e20, 6.0, 2.0]))
    ur = dev(np.array([5.0, 3.0, 1e20, 4.0]))
    y = dev(np.zeros(4))
    vp.op_dual_update(0.5, w, lr, ur, y)
    torch.cuda.synchronize()
    np.testing.assert_allclose(y.cpu().numpy(), [2.5, 2.5, -2.0, 0.0], rtol=1e-12, atol=1e-15)


def test_primal_update_kat():
    # test_cuda_kernels.cu:176-211: tau = 1 -> x {0,3,3}, x_bar {-1,4,3}
    import vajra_pdhg as vp
    v = dev(np.array([1.0, -2.0, 0.0]))
    c = dev(np.array([2.0, 1.0, 0.0]))
    lc = dev(np.array([0.0, -5.0, -1e20]))
    uc = dev(np.array([5.0, 5.0, 1e20]))
    x = dev(np.array([1.0, 2.0, 3.0]))
    xb = dev(np.zeros(3))
    vp.op_primal_update(1.0, v, c, lc, uc, x, xb)
    torch.cuda.synchronize()
    np.testing.assert_allclose(x.cpu().numpy(), [0.0, 3.0, 3.0], rtol=1e-12, atol=0)
    np.testing.assert_allclose(xb.cpu().numpy(), [-1.0, 4.0, 3.0], rtol=1e-12, atol=0)


def test_diagnostics_kat():
    # test_cuda_kernels.cu:214-251: primal residual 1.0, objective 8.0, no NaN
    import vajra_pdhg as vp
    out = vp.op_diagnostics(dev(np.array([5.0, 2.0])), dev(np.array([0.0, 3.0])), dev(np.array([4.0, 10.0])),
                            dev(np.zeros(2)), dev(np.array([2.0, 3.0])), dev(np.array([1.0, 2.0])), dev(np.zeros(2)),
                            dev(np.array([10.0, 10.0])))
    assert out[0] == 1.0 and out[2] == 8.0 and out[3] == 0.0


def test_complete_iteration_kat():
    # test_cuda_kernels.cu:254-359: one full iteration against the CPU formulas at 1e-12
    import vajra_pdhg as vp
    rp = np.array([0, 2, 4, 6], np.int64)
    ci = np.array([0, 1, 1, 2, 0, 2], np.int32)
    cv = np.array([1.0, 2.0, 1.0, -1.0, 2.0, 1.0])
    cp = np.array([0, 2, 4, 6], np.int64)
    ri = np.array([0, 2, 0, 1, 1, 2], np.int32)
    tv = np.array([1.0, 2.0, 2.0, 1.0, -1.0, 1.0])
    lr, ur = np.array([4.0, -1e20, 3.0]), np.array([4.0, 2.0, 1e20])
    c, lc, uc = np.array([1.0, -2.0, 3.0]), np.array([0.0, -1e20, -5.0]), np.array([10.0, 1e20, 5.0])
    x0, y0 = np.array([1.0, 0.0, 1.0]), np.zeros(3)
    sigma = tau = 0.4
    A = np.zeros((3, 3))
    for i in range(3):
        for p in range(rp[i], rp[i + 1]):
            A[i, ci[p]] = cv[p]
    w = A @ x0
    yh = y0 + sigma * w
    y1 = yh - sigma * np.clip(yh / sigma, lr, ur)
    v = A.T @ y1
    x1 = np.clip(x0 - tau * (v + c), lc, uc)
    xb1 = 2 * x1 - x0
    d_x, d_xb, d_y = dev(x0), dev(x0), dev(y0)
    d_w, d_v = torch.zeros(3, dtype=torch.float64, device="cuda"), torch.zeros(3, dtype=torch.float64, device="cuda")
    vp.op_spmv(dev(rp), dev(ci), dev(cv), d_xb, d_w, lanes=8)
    vp.op_dual_update(sigma, d_w, dev(lr), dev(ur), d_y)
    vp.op_spmv(dev(cp), dev(ri), dev(tv), d_y, d_v, lanes=8)
    vp.op_primal_update(tau, d_v, dev(c), dev(lc), dev(uc), d_x, d_xb)
    torch.cuda.synchronize()
    for got, want in [(d_w, w), (d_y, y1), (d_v, v), (d_x, x1), (d_xb, xb1)]:
        np.testing.assert_allclose(got.cpu().numpy(), want, rtol=1e-12, atol=1e-15)


def test_spmv_matches_oracle_cfg0():
    """Exact-order SpMV is bit-identical to the oracle's CSC products (sparse.cpp:196-238);
    lane-group SpMV agrees to a few ulp."""
    import vajra_pdhg as vp
    lp = lp_cases.cfg0_lp()
    rng = np.random.default_rng(7)
    x = rng.standard_normal(lp.n)
    x[rng.random(lp.n) < 0.05] = 1e-12          # exercised by the skip rule
    yv = rng.standard_normal(lp.m)
    rp, ci, av = lp.csr()
    want_y = O.csc_multiply(lp, x)
    want_v = O.csc_multiply_transpose(lp, yv)
    d_rp, d_ci, d_av, d_x = dev(rp), dev(ci), dev(av), dev(x)
    y = torch.zeros(lp.m, dtype=torch.float64, device="cuda")
    vp.op_spmv(d_rp, d_ci, d_av, d_x, y, exact_order=True, skip_tiny=True)
    v = torch.zeros(lp.n, dtype=torch.float64, device="cuda")
    vp.op_spmv(dev(lp.col_ptr), dev(lp.row_idx), dev(lp.vals), dev(yv), v, exact_order=True)
    torch.cuda.synchronize()
    assert np.array_equal(y.cpu().numpy(), want_y)
    assert np.array_equal(v.cpu().numpy(), want_v)
    dense_ref = lp.dense() @ x
    for lanes in (2, 4, 8, 16, 32, 64):
        y2 = torch.zeros(lp.m, dtype=torch.float64, device="cuda")
        vp.op_spmv(d_rp, d_ci, d_av, d_x, y2, lanes=lanes)
        torch.cuda.synchronize()
        np.testing.assert_allclose(y2.cpu().numpy(), dense_ref, rtol=1e-13, atol=1e-13)
