"""CPU tests of the oracle (oracle/pdhg_oracle.c), the comparison every GPU parity test relies on.

Pins, in order of strength:
  * golden vectors written by the reference compiled from its own sources
    (tests/golden/make_golden.py): trajectories, KAT answers, Ruiz factors, spectral norm;
  * the compiled reference itself (oracle/_ref), when it is present in this container;
  * closed forms and invariants that a dropped term, wrong sign or transposed operand breaks
    (dense brute force, diagonal / rank-one spectral norms, Ruiz norms, hand-computed KKT
    residuals, an optimal primal-dual pair from HiGHS).
"""
import json
import os

import numpy as np
import pytest

import lp_cases
import oracle as O
import vajra_pdhg as vp

GOLD = lp_cases.GOLDEN
META = json.load(open(os.path.join(GOLD, "golden_meta.json")))
TRAJ = [(1, 10**9), (10, 10**9), (100, 10**9), (1000, 10**9), (700, 100), (3000, 100)]


@pytest.fixture(scope="module")
def cfg0():
    return lp_cases.cfg0_lp()


@pytest.fixture(scope="module")
def gold_cfg0():
    return np.load(os.path.join(GOLD, "ref_cfg0.npz"))


@pytest.fixture(scope="module")
def gold_kats():
    return np.load(os.path.join(GOLD, "ref_kats.npz"))


def test_generator_matches_golden_inputs(cfg0):
    """The seeded generator still produces the LP the golden answers were computed on."""
    import hashlib
    h = hashlib.sha256()
    for a in (cfg0.col_ptr, cfg0.row_idx, cfg0.vals, cfg0.c, cfg0.col_lower, cfg0.col_upper, cfg0.row_lower,
              cfg0.row_upper):
        h.update(np.ascontiguousarray(a).tobytes())
    assert h.hexdigest() == META["cfg0_sha256"]


@pytest.mark.parametrize("iters,freq", TRAJ)
def test_oracle_trajectory_equals_reference(cfg0, gold_cfg0, iters, freq):
    o = O.pdhg_solve(cfg0, O.default_opts(max_iterations=iters, check_frequency=freq))
    g = lambda k: gold_cfg0[f"cfg0_{iters}_{freq}_{k}"]
    assert np.array_equal(o["x"], g("x"))
    assert np.array_equal(o["row_dual"], g("row_dual"))
    assert np.array_equal(o["col_dual"], g("col_dual"))
    for k in ("status", "iterations", "tau", "sigma", "restart_count", "primal_residual", "dual_residual",
              "duality_gap", "objective", "spectral_norm"):
        assert o[k] == g(k).item(), k


def test_oracle_ruiz_and_norm_equal_reference(cfg0, gold_cfg0):
    r = O.ruiz(cfg0)
    assert np.array_equal(r["row_scale"], gold_cfg0["cfg0_ruiz_row_scale"])
    assert np.array_equal(r["col_scale"], gold_cfg0["cfg0_ruiz_col_scale"])
    assert np.array_equal(r["vals"], gold_cfg0["cfg0_ruiz_vals"])
    assert np.array_equal(r["c"], gold_cfg0["cfg0_ruiz_c"])
    assert O.spectral_norm(cfg0, r["vals"]) == gold_cfg0["cfg0_spectral_norm_scaled"].item()


@pytest.mark.parametrize("name", sorted(lp_cases.kat_lps()))
def test_oracle_kats_equal_reference(gold_kats, name):
    lp, obj, xs = lp_cases.kat_lps()[name]
    o = O.pdhg_solve(lp, O.default_opts(max_iterations=20000, primal_tolerance=1e-5, dual_tolerance=1e-5))
    assert np.array_equal(o["x"], gold_kats[f"{name}_pdhg_x"])
    assert np.array_equal(o["row_dual"], gold_kats[f"{name}_pdhg_row_dual"])
    assert o["status"] == gold_kats[f"{name}_pdhg_status"].item()
    assert o["iterations"] == gold_kats[f"{name}_pdhg_iterations"].item()
    assert o["objective"] == gold_kats[f"{name}_pdhg_objective"].item()
    # the reference's dual simplex agrees with the published optimum
    assert gold_kats[f"{name}_simplex_status"].item() == 1
    assert abs(gold_kats[f"{name}_simplex_objective"].item() - obj) <= 1e-9 * (1 + abs(obj))


def test_oracle_infeasible_and_afiro_limit_equal_reference(gold_kats):
    o = O.pdhg_solve(lp_cases.infeasible_lp(), O.default_opts(max_iterations=500))
    assert o["status"] == gold_kats["infeasible_pdhg_status"].item() != 1
    assert np.array_equal(o["x"], gold_kats["infeasible_pdhg_x"])
    for name in ("afiro", "crude_blend"):
        lp = lp_cases.mps_lp(name)
        o = O.pdhg_solve(lp, O.default_opts(max_iterations=10))
        assert np.array_equal(o["x"], gold_kats[f"{name}_pdhg10_x"])
        assert o["status"] == gold_kats[f"{name}_pdhg10_status"].item()
        assert o["iterations"] == gold_kats[f"{name}_pdhg10_iterations"].item() == 11
        o = O.pdhg_solve(lp, O.default_opts(max_iterations=50000, primal_tolerance=1e-4, dual_tolerance=1e-4))
        assert np.array_equal(o["x"], gold_kats[f"{name}_pdhg_x"])
        assert o["iterations"] == gold_kats[f"{name}_pdhg_iterations"].item()
        assert abs(gold_kats[f"{name}_simplex_objective"].item() - lp_cases.MPS_OPTIMA[name]) <= 1e-6 * (
            1 + abs(lp_cases.MPS_OPTIMA[name]))


# ---- closed forms ----------------------------------------------------------------------
def random_lp(seed, m, n, density=0.3):
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((m, n)) * (rng.random((m, n)) < density)
    return vp.LP.from_dense(A, rng.standard_normal(n), -rng.random(n), 1 + rng.random(n), -np.ones(m), np.ones(m))


@pytest.mark.parametrize("seed", range(5))
def test_spmv_brute_force(seed):
    lp = random_lp(seed, 17, 23)
    rng = np.random.default_rng(100 + seed)
    x, y = rng.standard_normal(lp.n), rng.standard_normal(lp.m)
    x[::4] = 1e-12   # dropped by the reference's skip rule (sparse.cpp:207)
    A = lp.dense()
    xs = np.where(np.abs(x) <= 1e-11, 0.0, x)
    np.testing.assert_allclose(O.csc_multiply(lp, x), A @ xs, rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(O.csc_multiply_transpose(lp, y), A.T @ y, rtol=1e-13, atol=1e-13)
    rp, ci, v = O.csc_to_csr(lp)
    for i in range(lp.m):
        np.testing.assert_array_equal(ci[rp[i]:rp[i + 1]], np.nonzero(A[i])[0])
        np.testing.assert_array_equal(v[rp[i]:rp[i + 1]], A[i, np.nonzero(A[i])[0]])


def test_spmv_reference_kat():
    # tests/test_linalg.cpp / test_pdhg.cpp:442-477 style: 4x3 with an empty row
    A = np.array([[1.5, 0, -2.0], [0, 0, 0], [3.0, 4.0, -1.0], [0, 5.5, 0]])
    lp = vp.LP.from_dense(A, np.zeros(3), np.zeros(3), np.ones(3), np.zeros(4), np.ones(4))
    np.testing.assert_allclose(O.csc_multiply(lp, np.array([2.0, 3.0, 4.0])), [-5.0, 0.0, 14.0, 16.5])
    np.testing.assert_allclose(O.csc_multiply_transpose(lp, np.array([2.0, 4.0, 6.0, 1.0])), [21.0, 29.5, -10.0])


def test_spectral_norm_closed_forms():
    d = np.array([3.0, -7.0, 2.0, 0.5])
    lp = vp.LP.from_dense(np.diag(d), np.zeros(4), np.zeros(4), np.ones(4), np.zeros(4), np.ones(4))
    assert abs(O.spectral_norm(lp, max_iter=200, tol=1e-12) - 7.0) < 1e-6
    u, v = np.array([1.0, 2.0, -2.0]), np.array([0.5, 0.5, 0.5, 0.5])
    lp = vp.LP.from_dense(np.outer(u, v), np.zeros(4), np.zeros(4), np.ones(4), np.zeros(3), np.ones(3))
    assert abs(O.spectral_norm(lp) - np.linalg.norm(u) * np.linalg.norm(v)) < 1e-12
    for seed in range(3):
        lp = random_lp(seed, 30, 40)
        smax = np.linalg.svd(lp.dense(), compute_uv=False)[0]
        est = O.spectral_norm(lp, max_iter=500, tol=1e-14)
        assert est <= smax * (1 + 1e-9) and est > 0.999 * smax
    empty = vp.LP(2, 2, np.zeros(3, np.int64), np.zeros(0, np.int32), np.zeros(0), np.zeros(2), np.zeros(2),
                  np.ones(2), np.zeros(2), np.ones(2))
    assert O.spectral_norm(empty) == 1.0


def test_ruiz_invariants():
    # test_linalg.cpp:283-311: after scaling, row and column inf-norms are close to 1
    lp = random_lp(3, 40, 60)
    lp.vals *= np.exp(np.random.default_rng(0).uniform(-6, 6, lp.nnz))
    r = O.ruiz(lp, max_iters=10, tol=0.05)
    A = vp.LP(lp.m, lp.n, lp.col_ptr, lp.row_idx, r["vals"], lp.c, lp.col_lower, lp.col_upper, lp.row_lower,
              lp.row_upper).dense()
    rows, cols = np.abs(A).max(1), np.abs(A).max(0)
    assert np.all(np.abs(rows[rows > 0] - 1) < 0.25) and np.all(np.abs(cols[cols > 0] - 1) < 0.25)
    # the scaled matrix is R A C with the returned factors
    np.testing.assert_allclose(A, np.diag(r["row_scale"]) @ lp.dense() @ np.diag(r["col_scale"]), rtol=1e-12)


def test_rel_kkt_hand_computed():
    # min x1 + x2  s.t.  x1 + x2 >= 2 (row 0), x1 - x2 = 0 (row 1), 0 <= x <= 5
    lp = vp.LP.from_dense([[1, 1], [1, -1]], [1, 1], [0, 0], [5, 5], [2, 0], [1e20, 0])
    # optimal pair: x = (1,1), y = (1, 0): d = c - A^T y = 0, dual obj = 2 = primal obj
    k = O.rel_kkt(lp, np.array([1.0, 1.0]), np.array([1.0, 0.0]))
    assert k[0] < 1e-15 and k[4] == 2.0 and k[5] == 2.0
    # primal violation: x = (0.5, 0.5) violates row 0 by 1 -> rel_p = 1 / (1 + |b|), b = (2, 0)
    k = O.rel_kkt(lp, np.array([0.5, 0.5]), np.array([1.0, 0.0]))
    assert abs(k[1] - 1.0 / 3.0) < 1e-15
    # dual sign violation on the >= row: y0 = -1 has no finite upper bound to price
    k = O.rel_kkt(lp, np.array([1.0, 1.0]), np.array([-1.0, 0.0]))
    # d = (2, 2) priced by lower bounds 0 -> lambda = d; residual only from y0: |y0| / (1 + |c|)
    assert abs(k[2] - 1.0 / (1 + np.sqrt(2))) < 1e-15
    # transposed operand (A^T y computed as A y) would break this: y = (0, 1) -> d = (0, 2)
    k = O.rel_kkt(lp, np.array([1.0, 1.0]), np.array([0.0, 1.0]))
    assert abs(k[5] - (0 + 0)) < 1e-15 and abs(k[3] - 2.0 / 3.0) < 1e-15


def test_rel_kkt_of_highs_optimum_is_small():
    """An optimal primal-dual pair from a library solver has (near) zero rel-KKT; for equality
    rows HiGHS's eqlin marginals are d(objective)/d(b), i.e. the canonical dual."""
    pytest.importorskip("scipy")
    from scipy.optimize import linprog
    lp = random_lp(11, 20, 35)
    A = lp.dense()
    x0 = np.random.default_rng(5).uniform(lp.col_lower, lp.col_upper)
    lp.row_lower = lp.row_upper = A @ x0
    r = linprog(lp.c, A_eq=A, b_eq=lp.row_lower, bounds=list(zip(lp.col_lower, lp.col_upper)), method="highs")
    assert r.status == 0
    k = O.rel_kkt(lp, r.x, r.eqlin.marginals)
    assert k[0] < 1e-8, k
    assert O.rel_kkt(lp, r.x, -r.eqlin.marginals)[0] > 1e-3   # a sign error is caught


def test_quality_and_verifier_known_violations():
    lp = vp.LP.from_dense([[1, 1], [1, 0]], [1, 2], [0, 0], [4, 4], [1, -1e20], [1e20, 3])
    # x1 = 3.5 violates row 1 (<= 3) by 0.5; col bound fine
    x = np.array([3.5, 0.0])
    q = O.recompute_quality(lp, x, np.array([1.0, 0.0]))
    assert abs(q["max_primal_violation"] - 0.5) < 1e-15 and not q["is_primal_feasible"]
    v = O.verify_solution(lp, x, np.array([1.0, 0.0]), 3.5, tol=1e-6)
    assert not v["passed"] and abs(v["max_primal_violation"] - 0.5) < 1e-15 and v["objective_matches"]
    v = O.verify_solution(lp, np.array([1.0, 0.0]), np.array([1.0, 0.0]), 1.0, tol=1e-6)
    assert v["passed"], v


@pytest.mark.skipif(not O.have_ref(), reason="oracle/_ref not built (no reference checkout)")
@pytest.mark.parametrize("seed", range(4))
def test_oracle_equals_compiled_reference(seed):
    rng = np.random.default_rng(seed)
    lp = random_lp(seed, 25, 40)
    lp.row_lower = rng.uniform(-2, 0, lp.m)
    lp.row_upper = lp.row_lower + rng.uniform(0, 3, lp.m)
    lp.row_upper[::5] = 1e20
    lp.col_lower[::7] = -1e20
    lp.maximize = bool(seed % 2)
    for iters, freq in ((37, 10), (500, 100), (2000, 50)):
        o = O.pdhg_solve(lp, O.default_opts(max_iterations=iters, check_frequency=freq))
        r = O.ref_solve(lp, iteration_limit=iters, check_frequency=freq)
        assert np.array_equal(o["x"], r["x"]) and np.array_equal(o["row_dual"], r["row_dual"])
        assert np.array_equal(o["col_dual"], r["col_dual"]) and np.array_equal(o["row_value"], r["row_value"])
        assert o["status"] == r["status"] and o["iterations"] == r["iterations"]
        assert o["tau"] == r["tau"] and o["restart_count"] == r["restart_count"]
    x = rng.standard_normal(lp.n)
    yd = rng.standard_normal(lp.m)
    q = O.recompute_quality(lp, x, yd)
    qr = np.zeros(5)
    s, keep = O._lp_struct(lp)
    O.ref_lib().ref_quality(lp.m, lp.n, O._p(lp.col_ptr), O._p(lp.row_idx), O._p(lp.vals), O._p(lp.c),
                            O._p(lp.col_lower), O._p(lp.col_upper), O._p(lp.row_lower), O._p(lp.row_upper),
                            int(lp.maximize), O._c_dbl(lp.offset), O._p(x), O._p(yd), None, O._p(qr))
    assert [q["max_primal_violation"], q["max_dual_violation"], q["max_complementarity_violation"]] == list(qr[:3])
    vr = np.zeros(10)
    O.ref_lib().ref_verify(lp.m, lp.n, O._p(lp.col_ptr), O._p(lp.row_idx), O._p(lp.vals), O._p(lp.c),
                           O._p(lp.col_lower), O._p(lp.col_upper), O._p(lp.row_lower), O._p(lp.row_upper),
                           int(lp.maximize), O._c_dbl(lp.offset), O._p(x), O._p(yd), O._c_dbl(1.5), 1, O._c_dbl(1e-6),
                           O._p(vr))
    v = O.verify_solution(lp, x, yd, 1.5, tol=1e-6)
    assert [v["passed"], v["max_primal_violation"], v["max_dual_violation"], v["objective_discrepancy"]] == [
        bool(vr[0]), vr[1], vr[2], vr[4]]


def test_mps_reader_against_published_optima():
    meta = META
    for name in ("afiro", "crude_blend"):
        lp = lp_cases.mps_lp(name)
        assert abs(meta[f"{name}_highs_objective"] - lp_cases.MPS_OPTIMA[name]) < 1e-6 * (1 + abs(lp_cases.MPS_OPTIMA[name]))
    afiro = lp_cases.mps_lp("afiro")
    assert (afiro.m, afiro.n, afiro.nnz) == (27, 32, 83)


@pytest.mark.skipif(not O.have_ref(), reason="oracle/_ref not built (no reference checkout)")
@pytest.mark.parametrize("seed,m,n", [(0, 12, 30), (1, 30, 12), (2, 1, 9), (3, 9, 1), (4, 40, 40), (5, 3, 3)])
def test_oracle_equals_compiled_reference_structured(seed, m, n):
    """The restated CPU PDHG (oracle) against the reference's solve_pdhg_cpu compiled from its
    sources, bit for bit (iterates, duals, row values, status, iteration count, step sizes,
    restarts), on structured LPs and at several check frequencies, including the limit
    reached between checks."""
    lp = lp_cases.structured_lp(seed, m, n)
    for iters, freq in ((1, 1), (23, 7), (400, 100), (1500, 40)):
        o = O.pdhg_solve(lp, O.default_opts(max_iterations=iters, check_frequency=freq))
        r = O.ref_solve(lp, iteration_limit=iters, check_frequency=freq)
        assert o["status"] == r["status"] and o["iterations"] == r["iterations"], (iters, freq, o["status"], r["status"])
        assert np.array_equal(o["x"], r["x"]) and np.array_equal(o["row_dual"], r["row_dual"])
        assert np.array_equal(o["col_dual"], r["col_dual"]) and np.array_equal(o["row_value"], r["row_value"])
        assert o["tau"] == r["tau"] and o["sigma"] == r["sigma"] and o["restart_count"] == r["restart_count"]
        assert o["objective"] == r["objective"]
    # quality audit and verifier at a random point (model.cpp:203-374, verifier.cpp:25-330)
    rng = np.random.default_rng(50 + seed)
    x, yd = rng.standard_normal(lp.n), rng.standard_normal(lp.m)
    q = O.recompute_quality(lp, x, yd)
    qr = np.zeros(5)
    O.ref_lib().ref_quality(lp.m, lp.n, O._p(lp.col_ptr), O._p(lp.row_idx), O._p(lp.vals), O._p(lp.c),
                            O._p(lp.col_lower), O._p(lp.col_upper), O._p(lp.row_lower), O._p(lp.row_upper),
                            int(lp.maximize), O._c_dbl(lp.offset), O._p(x), O._p(yd), None, O._p(qr))
    assert [q["max_primal_violation"], q["max_dual_violation"], q["max_complementarity_violation"]] == list(qr[:3])
    vr = np.zeros(10)
    O.ref_lib().ref_verify(lp.m, lp.n, O._p(lp.col_ptr), O._p(lp.row_idx), O._p(lp.vals), O._p(lp.c),
                           O._p(lp.col_lower), O._p(lp.col_upper), O._p(lp.row_lower), O._p(lp.row_upper),
                           int(lp.maximize), O._c_dbl(lp.offset), O._p(x), O._p(yd), O._c_dbl(0.75), 1, O._c_dbl(1e-6),
                           O._p(vr))
    v = O.verify_solution(lp, x, yd, 0.75, tol=1e-6)
    assert [v["passed"], v["max_primal_violation"], v["max_dual_violation"], v["objective_discrepancy"]] == [
        bool(vr[0]), vr[1], vr[2], vr[4]]


@pytest.mark.skipif(not O.have_ref(), reason="oracle/_ref not built (no reference checkout)")
@pytest.mark.parametrize("m,n", [(0, 4), (3, 0)])
def test_oracle_trivial_sizes_equal_compiled_reference(m, n):
    """LPs without rows or without columns take the reference's trivial exits
    (pdhg.cpp:88-127); the oracle returns the same status, point and objective."""
    rng = np.random.default_rng(7)
    A = np.zeros((m, n))
    lp = vp.LP.from_dense(A, rng.standard_normal(n), -rng.random(n), 1 + rng.random(n), -np.ones(m), np.ones(m))
    o = O.pdhg_solve(lp, O.default_opts(max_iterations=100, check_frequency=10))
    r = O.ref_solve(lp, iteration_limit=100, check_frequency=10)
    assert o["status"] == r["status"] and o["iterations"] == r["iterations"]
    assert np.array_equal(o["x"], r["x"]) and np.array_equal(o["row_dual"], r["row_dual"])
    assert o["objective"] == r["objective"]
