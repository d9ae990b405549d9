"""Device-side setup of the restarted engine, pinned:

* the CSR of A built on the device from the uploaded CSC (stable sort by row, the order of
  SparseMatrixCSC::to_csr, src/linalg/sparse.cpp:147-184) and the kernel layouts built from it
  on the device (row tiles, XCD column split, sliced ELL): solves from a device-built CSR are
  bit-identical to solves from the host-built CSR, for every kernel variant, single and batched;
* the device power iteration (the engine's counterpart of estimate_spectral_norm,
  src/solvers/pdhg/pdhg.cpp:11-56) against closed forms: a diagonal matrix gives max |a_ii|,
  a rank-one matrix u v^T gives |u| |v|; on random sparse matrices it never exceeds the
  largest singular value (numpy SVD) and comes within 1% of it;
* the device Ruiz passes (the engine's counterpart of RuizScaler::compute_and_scale,
  src/linalg/scaling.cpp:7-153) against the oracle's restatement of the reference's Ruiz,
  pass for pass (tolerance 0, so the reference runs every pass), on LPs whose factors stay
  inside the device's 1e+-6 bound."""
import numpy as np
import pytest

import lp_cases
import oracle as O
import vajra_pdhg as vp

pytestmark = pytest.mark.gpu

# forced kernel choices: two uploads autotune independently, and near-tied candidates may
# time apart differently between them
VARIANTS = [dict(lanes_A=-2, lanes_At=-2, col_splits_A=1, col_splits_At=1),
            dict(lanes_A=-1, lanes_At=-1, col_splits_A=1, col_splits_At=1),
            dict(lanes_A=-3, lanes_At=-3), dict(col_splits_A=4, col_splits_At=2, lanes_A=-1, lanes_At=-1),
            dict(col_splits_A=8, col_splits_At=2, lanes_A=2, lanes_At=4)]


def solve_point(engine, lp, host_csr, **kw):
    engine.upload(lp, host_csr=host_csr)
    r = engine.solve(vp.default_params(vp.MODE_RESTARTED, tolerance=1e-6, max_iterations=20000, **kw))
    x, _, y = engine.download()
    return r, x, y


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_device_csr_and_layouts_equal_host(engine, prec):
    lps = [lp_cases.cfg0_lp(), lp_cases.structured_feasible_lp(5, 60, 150), lp_cases.structured_feasible_lp(6, 150, 60),
           vp.gen_random_lp(0x5EED2026 + 21, 40_000, 100_000, 6, 0.1, 3.0)]
    for lp in lps:
        for kw in VARIANTS:
            print(f"[case] device vs host CSR {prec} {lp.m}x{lp.n} {kw}", flush=True)
            a = solve_point(engine, lp, True, precision=prec, **kw)
            b = solve_point(engine, lp, False, precision=prec, **kw)
            assert (a[0].iterations, a[0].status, a[0].lanes_A, a[0].lanes_At) == \
                   (b[0].iterations, b[0].status, b[0].lanes_A, b[0].lanes_At)
            assert np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2])


def test_device_csr_batch_equals_host(engine):
    lps = [lp_cases.structured_feasible_lp(*s) for s in [(0, 12, 30), (1, 30, 12), (5, 60, 150), (2, 1, 9)]]
    lps += [vp.gen_random_lp(1000 + i, m, int(2.5 * m)) for i, m in enumerate([500, 1300])]
    out = []
    for host_csr in (True, False):
        engine.upload_batch(lps, host_csr=host_csr)
        res = engine.solve_batch(vp.default_params(vp.MODE_RESTARTED, tolerance=1e-6, max_iterations=50000))
        xs, ys = engine.download_batch()
        out.append(([r.iterations for r in res], xs, ys))
    assert out[0][0] == out[1][0]
    for a, b in zip(out[0][1] + out[0][2], out[1][1] + out[1][2]):
        assert np.array_equal(a, b)


def spectral(engine, lp):
    """The device power-iteration estimate of |A|_2 with scaling off (A as given)."""
    engine.upload(lp)
    r = engine.solve(vp.default_params(vp.MODE_RESTARTED, tolerance=0.0, max_iterations=64, check_frequency=64,
                                       enable_scaling=0, bound_objective_rescaling=0))
    return r.spectral_norm


def dense_lp(A):
    m, n = A.shape
    return vp.LP.from_dense(A, np.ones(n), np.zeros(n), np.full(n, 10.0), -np.ones(m), np.ones(m))


def test_power_iteration_closed_forms(engine):
    rng = np.random.default_rng(3)
    d = rng.uniform(0.5, 2.0, 300)
    d[137] = -5.0
    got = spectral(engine, dense_lp(np.diag(d)))
    assert abs(got - 5.0) <= 1e-12 * 5.0, got
    u, v = rng.standard_normal(40), rng.standard_normal(60)
    want = np.linalg.norm(u) * np.linalg.norm(v)
    got = spectral(engine, dense_lp(np.outer(u, v)))
    assert abs(got - want) <= 1e-12 * want, (got, want)
    for seed in range(4):
        A = rng.standard_normal((150, 220)) * (rng.random((150, 220)) < 0.1)
        smax = np.linalg.svd(A, compute_uv=False)[0]
        got = spectral(engine, dense_lp(A))
        assert got <= smax * (1 + 1e-12) and got >= 0.99 * smax, (seed, got, smax)


@pytest.mark.parametrize("passes", [1, 4, 10])
def test_device_ruiz_equals_reference_ruiz(engine, passes):
    for lp in [lp_cases.cfg0_lp(), vp.gen_random_lp(0x5EED2026 + 22, 3000, 7500, 6, 0.1, 3.0)]:
        print(f"[case] device Ruiz {passes} passes {lp.m}x{lp.n}", flush=True)
        engine.upload(lp)
        engine.solve(vp.default_params(vp.MODE_RESTARTED, tolerance=0.0, max_iterations=64, check_frequency=64,
                                       ruiz_iterations=passes, pock_chambolle=0))
        rs, cs = engine.scaling_factors(lp.m, lp.n)
        o = O.ruiz(lp, max_iters=passes, tol=0.0)
        assert np.all((o["row_scale"] > 1e-6) & (o["row_scale"] < 1e6))   # the device bound is inactive
        assert np.all((o["col_scale"] > 1e-6) & (o["col_scale"] < 1e6))
        np.testing.assert_allclose(rs, o["row_scale"], rtol=1e-13, atol=0)
        np.testing.assert_allclose(cs, o["col_scale"], rtol=1e-13, atol=0)
