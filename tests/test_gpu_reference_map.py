"""Reference iteration map on the GPU (VJ_MODE_REFERENCE) against the oracle, bit for bit.

The oracle restates CpuPdhgSolver::solve (pdhg.cpp:70-440) and is itself pinned bit-exactly
to the reference compiled from its sources (tests/test_oracle.py and the golden files), so
equality here means the GPU reproduces the reference CPU PDHG: the same iterates, step
sizes, rebalancing, restarts, best iterate, status and iteration count."""
import numpy as np
import pytest

import lp_cases
import oracle as O
import vajra_pdhg as vp

pytestmark = pytest.mark.gpu


def gpu_ref_solve(engine, lp, **kw):
    engine.upload(lp)
    p = vp.default_params(vp.MODE_REFERENCE)
    for k, v in kw.items():
        setattr(p, k, v)
    r = engine.solve(p)
    x, rd, y = engine.download()
    return r, x, rd, y


def oracle_solve(lp, **kw):
    o = O.default_opts()
    for k, v in kw.items():
        setattr(o, k, v)
    return O.pdhg_solve(lp, o)


@pytest.mark.parametrize("iters,freq", [(1, 10**9), (10, 10**9), (100, 10**9), (1000, 10**9), (700, 100), (3000, 100)])
def test_cfg0_trajectory_bit_exact(engine, iters, freq):
    """Plain Chambolle-Pock iterates (no checks: freq huge -> one check at the last iteration)
    and the full reference loop with checks every 100 iterations."""
    lp = lp_cases.cfg0_lp()
    print(f"[case] cfg0 iters={iters} freq={freq}", flush=True)
    r, x, rd, _ = gpu_ref_solve(engine, lp, max_iterations=iters, check_frequency=freq)
    o = oracle_solve(lp, max_iterations=iters, check_frequency=freq)
    assert np.array_equal(x, o["x"])
    assert np.array_equal(rd, o["row_dual"])
    assert r.status == o["status"]
    assert r.iterations == o["iterations"]
    assert r.restart_count == o["restart_count"]
    assert r.tau == o["final_tau"] and r.sigma == o["final_sigma"]
    assert r.primal_residual == o["primal_residual"]
    assert r.dual_residual == o["dual_residual"]
    assert r.duality_gap == o["duality_gap"]


def test_cfg0_scaled_iterates_bit_exact(engine):
    lp = lp_cases.cfg0_lp()
    engine.upload(lp)
    engine.set_params(vp.default_params(vp.MODE_REFERENCE, check_frequency=10**9, max_iterations=10**6))
    for k in (1, 37, 250):
        x, xb, y = engine.reference_iterates(k)
        o = O.pdhg_solve(lp, O.default_opts(stop_at_iter=k, check_frequency=10**9), want_traj=True)
        assert np.array_equal(x, o["traj_x"]) and np.array_equal(xb, o["traj_xbar"]) and np.array_equal(y, o["traj_y"])


@pytest.mark.parametrize("name", sorted(lp_cases.kat_lps()))
def test_reference_kats_bit_exact(engine, name):
    """The reference's PDHG known-answer LPs (test_pdhg.cpp:54-436) at tolerance 1e-5,
    20,000 iterations: identical outcome to the oracle, and the published optimum to 1e-3."""
    lp, obj, xs = lp_cases.kat_lps()[name]
    print(f"[case] kat {name}", flush=True)
    r, x, rd, _ = gpu_ref_solve(engine, lp, max_iterations=20000, tolerance=1e-5)
    o = oracle_solve(lp, max_iterations=20000, primal_tolerance=1e-5, dual_tolerance=1e-5)
    assert np.array_equal(x, o["x"]) and np.array_equal(rd, o["row_dual"])
    assert r.status == o["status"] and r.iterations == o["iterations"]
    got_obj = float(np.dot(lp.c, x)) + lp.offset
    assert abs(got_obj - o["objective"]) <= 1e-12 * (1 + abs(obj))
    if r.status == 1:
        assert abs(got_obj - obj) < 1e-3
        if xs is not None:
            np.testing.assert_allclose(x, xs, atol=1e-3)


def test_reference_infeasible_not_optimal(engine):
    # test_pdhg.cpp:317-348: never claims OPTIMAL on an infeasible LP
    lp = lp_cases.infeasible_lp()
    r, x, rd, _ = gpu_ref_solve(engine, lp, max_iterations=500)
    o = oracle_solve(lp, max_iterations=500)
    assert r.status != 1 and r.status == o["status"]
    assert np.array_equal(x, o["x"])


def test_reference_afiro_iteration_limit(engine):
    # test_pdhg.cpp:351-365: iteration_limit 10 -> ITERATION_LIMIT or FEASIBLE, never OPTIMAL
    lp = lp_cases.mps_lp("afiro")
    r, x, rd, _ = gpu_ref_solve(engine, lp, max_iterations=10)
    o = oracle_solve(lp, max_iterations=10)
    assert r.status in (5, 6) and r.status == o["status"] and r.iterations == o["iterations"] == 11
    assert np.array_equal(x, o["x"])


@pytest.mark.parametrize("seed,m,n", [(0, 12, 30), (1, 30, 12), (2, 1, 9), (3, 9, 1), (4, 40, 40), (5, 3, 3)])
def test_structured_lps_bit_exact(engine, seed, m, n):
    """LPs with empty rows and columns, fixed, free and one-sided columns, equality and
    one-sided rows, tiny coefficients, offsets and both senses (the same LPs the oracle is
    pinned on against the compiled reference): the device reference map equals the oracle
    bit for bit, at several check frequencies, including a limit reached between checks."""
    lp = lp_cases.structured_lp(seed, m, n)
    for iters, freq in ((1, 1), (23, 7), (400, 100), (1500, 40)):
        print(f"[case] structured seed={seed} {m}x{n} iters={iters} freq={freq}", flush=True)
        r, x, rd, _ = gpu_ref_solve(engine, lp, max_iterations=iters, check_frequency=freq)
        o = oracle_solve(lp, max_iterations=iters, check_frequency=freq)
        assert np.array_equal(x, o["x"]) and np.array_equal(rd, o["row_dual"])
        assert r.status == o["status"] and r.iterations == o["iterations"]
        assert r.restart_count == o["restart_count"]
        assert r.tau == o["final_tau"] and r.sigma == o["final_sigma"]
