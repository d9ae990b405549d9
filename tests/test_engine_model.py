"""The numpy model of the restarted engine (tests/engine_model.py) on the LP families that
found the equilibration stalls: without the bound on the cumulative equilibration factors,
and without leaving the bounded rows and columns out of the rescaling norms, the iteration
stalls (rel-KKT frozen far above the tolerance); with both, every LP of the families
converges and meets its SciPy-HiGHS objective. The device tests
(tests/test_gpu_restarted.py) check the same LPs on the engine itself."""
import numpy as np
import pytest

import engine_model as M
import lp_cases


def _lp(seed, m, n, dec):
    lp = lp_cases.structured_feasible_lp(seed, m, n)
    return lp_cases.rescaled_lp(lp, seed, dec) if dec else lp


def test_unbounded_equilibration_stalls():
    """A row and a column of tiny coefficients only: equilibrated without a bound, their
    bounds / costs grow by up to 4e10 and the rescaling crushes everything else."""
    lp = _lp(0, 12, 30, 1.0)
    r = M.solve(lp, max_iters=20000, max_scale=None)
    assert r["status"] == 6 and r["kkt"][0] > 0.5, r["kkt"]
    r = M.solve(lp, max_iters=20000)
    assert r["status"] == 1, r["kkt"]


def test_near_empty_column_needs_the_norm_exclusion():
    """A column with one tiny coefficient and a real cost, scaled to the bound: its inflated
    cost still dominates the objective norm unless bounded columns are left out of it."""
    lp = _lp(11, 2, 64, 2.0)
    r = M.solve(lp, max_iters=40000, exclude_bounded=False)
    assert r["status"] == 6, r["kkt"]
    r = M.solve(lp, max_iters=40000)
    assert r["status"] == 1, r["kkt"]


@pytest.mark.parametrize("dec", [0.0, 1.0, 3.0])
def test_structured_and_rescaled_lps_converge(dec):
    for s, m, n in [(0, 12, 30), (1, 30, 12), (2, 1, 9), (3, 9, 1), (4, 40, 40), (5, 60, 150), (6, 150, 60)]:
        lp = _lp(s, m, n, dec)
        opt = lp_cases.highs_optimum(lp)
        r = M.solve(lp, max_iters=60000)
        assert r["status"] == 1, (s, dec, r["kkt"][:4])
        got = float(np.dot(lp.c, r["x"])) + lp.offset
        assert abs(got - opt) <= 1e-5 * (1 + abs(opt)), (s, dec, got, opt)


@pytest.mark.parametrize("rho", [1.0, 0.5])
def test_refine_every_check_changes_nothing_in_exact_arithmetic(rho):
    """Moving the centre at a check and shifting the iterate and the Halpern anchor with it is a
    change of coordinates: in exact arithmetic the iteration is the same. In FP64 the runs with
    and without it must then agree to rounding over blocks that cross restarts and re-centrings
    (a wrong sign or a missing anchor term in the shift gives O(1) differences), and the restart
    decisions must be the same."""
    for lp in (lp_cases.cfg0_lp(), lp_cases.structured_feasible_lp(5, 60, 150)):
        a = M.solve(lp, K=64, tol=0.0, max_iters=12 * 64, rho=rho)
        b = M.solve(lp, K=64, tol=0.0, max_iters=12 * 64, rho=rho, refine_every_check=True)
        assert a["restarts"] == b["restarts"] and 1 <= a["restarts"] < 12, (a["restarts"], b["restarts"])
        for v in ("x", "y"):
            scale = max(1.0, float(np.max(np.abs(a[v]))))
            assert float(np.max(np.abs(a[v] - b[v]))) <= 1e-9 * scale, v


def test_refine_every_check_reaches_the_optimum():
    lp = lp_cases.structured_feasible_lp(6, 150, 60)
    opt = lp_cases.highs_optimum(lp)
    r = M.solve(lp, max_iters=60000, refine_every_check=True)
    assert r["status"] == 1, r["kkt"][:4]
    got = float(np.dot(lp.c, r["x"])) + lp.offset
    assert abs(got - opt) <= 1e-5 * (1 + abs(opt)), (got, opt)


def test_bf16_rounding_is_nearest_even():
    """The model's BF16 rounding against a brute-force choice between the two neighbouring
    BF16 values (nearest, ties to the even significand)."""
    rng = np.random.default_rng(7)
    v = np.concatenate([rng.standard_normal(5000).astype(np.float32) * np.float32(10.0) ** rng.integers(-8, 8, 5000),
                        np.array([1.0, -1.0, 1 + 2 ** -8, 1 + 3 * 2 ** -8, 0.0, 65504.0], np.float32)]).astype(np.float32)
    got = M._bf16(v)
    bits = v.view(np.uint32)
    lo = (bits & np.uint32(0xFFFF0000)).view(np.float32)
    hi = ((bits & np.uint32(0xFFFF0000)) + np.uint32(0x10000)).view(np.float32)
    dl, dh = np.abs(v.astype(float) - lo.astype(float)), np.abs(hi.astype(float) - v.astype(float))
    even_lo = ((lo.view(np.uint32) >> 16) & 1) == 0
    want = np.where(dl < dh, lo, np.where(dh < dl, hi, np.where(even_lo, lo, hi)))
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


@pytest.mark.parametrize("values,refine", [("fp32", False), ("bf16", True), ("fp16", True)])
def test_rounded_sweeps_reach_1e6(values, refine):
    """The low-precision model (values rounded, iterates in FP32, FP64 centre and checks)
    reaches rel-KKT 1e-6 on the 2,000 x 5,000 LP (with at most 3x the FP64 model's iterations,
    a sanity bound against a near-stall: BF16 takes about 1.7x)."""
    lp = lp_cases.cfg0_lp()
    S = M.setup(lp)
    ref = M.solve(lp, S=S, max_iters=100000)
    r = M.solve_rounded(lp, S=S, values=values, refine_every_check=refine, max_iters=100000)
    assert r["status"] == 1 and r["kkt"] <= 1e-6, r
    assert r["iters"] <= 3 * ref["iters"], (r["iters"], ref["iters"])
