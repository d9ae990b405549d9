"""The production fused kernels of the restarted engine, pinned step by step: FP64 solves
stopped after exactly B check blocks of K = 64 iterations (tolerance 0, max_iterations = B K)
return the full iterate of the last check (centre + the last T(z) correction, unscaled), and
that point must equal the numpy model of the engine (tests/engine_model.py: the same setup,
correction coordinates, reflected Halpern step, restart rule and primal-weight update, written
with scipy sparse products) to 1e-12 relative, for every kernel variant the autotuner can pick:
tile-streamed, sliced ELL, lane groups, XCD column splits of A and of A^T, LDS-staged, and the
batched (block-diagonal) kernels. B = 1 checks the fused steps from the initial centre; B = 3 also
crosses the restart (re-centring and the shifted data) that the first check always takes.

The differences come only from summation order (scipy's products against the kernels' fixed
orders) and from the setup's rounding, so they stay near 1e-15 relative."""
import numpy as np
import pytest

import engine_model as M
import lp_cases
import vajra_pdhg as vp

pytestmark = pytest.mark.gpu

K = 64
VARIANTS = {
    "tile_streamed": dict(lanes_A=-1, lanes_At=-1, col_splits_A=1, col_splits_At=1),
    "sliced_ell": dict(lanes_A=-3, lanes_At=-3, col_splits_A=1, col_splits_At=1),
    "lane_groups_8": dict(lanes_A=8, lanes_At=8, col_splits_A=1, col_splits_At=1),
    "split_A4_At2_streamed": dict(lanes_A=-1, lanes_At=-1, col_splits_A=4, col_splits_At=2),
    "split_A2_At4_groups": dict(lanes_A=2, lanes_At=4, col_splits_A=2, col_splits_At=4),
    "lds_staged": dict(lanes_A=-5, lanes_At=-5, col_splits_A=1, col_splits_At=1),
    "lds_staged_w16": dict(lanes_A=-7, lanes_At=-7, col_splits_A=1, col_splits_At=1),
    "lds_staged_256": dict(lanes_A=-5, lanes_At=-5, col_splits_A=1, col_splits_At=1, stage_threads=256),
    "lds_staged_w16_256": dict(lanes_A=-7, lanes_At=-7, col_splits_A=1, col_splits_At=1, stage_threads=256),
}


def model_point(lp, blocks, rho=1.0, **kw):
    r = M.solve(lp, K=K, tol=0.0, max_iters=blocks * K, rho=rho, **kw)
    assert r["iters"] == blocks * K
    return r["x"], r["y"], r["restarts"]


def close(got, want, what):
    scale = max(1.0, float(np.max(np.abs(want))) if want.size else 1.0)
    err = float(np.max(np.abs(got - want))) / scale if want.size else 0.0
    assert err <= 1e-12, (what, err)
    return err


def lps():
    return [("cfg0", lp_cases.cfg0_lp()), ("structured_5", lp_cases.structured_feasible_lp(5, 60, 150)),
            ("structured_6", lp_cases.structured_feasible_lp(6, 150, 60))]


@pytest.mark.parametrize("blocks", [1, 3])
@pytest.mark.parametrize("variant", sorted(VARIANTS))
def test_single_lp_blocks_equal_model(engine, variant, blocks):
    worst = 0.0
    for name, lp in lps():
        print(f"[case] step parity {variant} blocks={blocks} {name}", flush=True)
        engine.upload(lp)
        r = engine.solve(vp.default_params(vp.MODE_RESTARTED, precision="fp64", tolerance=0.0,
                                           max_iterations=blocks * K, check_frequency=K, **VARIANTS[variant]))
        x, _, y = engine.download()
        assert r.iterations == blocks * K and r.checks == blocks, r.as_dict()
        mx, my, restarts = model_point(lp, blocks)
        assert r.restart_count == restarts, (r.restart_count, restarts)
        worst = max(worst, close(x, mx, (name, "x")), close(y, my, (name, "y")))
    print(f"[case] worst relative difference {worst:.2e}", flush=True)


@pytest.mark.parametrize("blocks", [1, 3])
@pytest.mark.parametrize("lanes,threads", [(-1, 0), (8, 0), (-5, 1024), (-7, 1024), (-5, 256), (-7, 256)])
def test_batch_blocks_equal_model(engine, lanes, threads, blocks):
    """The batched kernels (every instance in one launch, per-instance state) on a batch of
    LPs of different shapes: each instance equals its own model run (the staged kernels in
    both workgroup shapes)."""
    cases = lps()
    engine.upload_batch([lp for _, lp in cases])
    print(f"[case] batch step parity lanes={lanes} threads={threads} blocks={blocks}", flush=True)
    res = engine.solve_batch(vp.default_params(vp.MODE_RESTARTED, precision="fp64", tolerance=0.0,
                                               max_iterations=blocks * K, check_frequency=K, lanes_A=lanes,
                                               lanes_At=lanes, stage_threads=threads))
    xs, ys = engine.download_batch()
    for (name, lp), r, x, y in zip(cases, res, xs, ys):
        assert r.iterations == blocks * K, r.as_dict()
        mx, my, restarts = model_point(lp, blocks)
        assert r.restart_count == restarts
        close(x, mx, (name, "x"))
        close(y, my, (name, "y"))


@pytest.mark.parametrize("variant", ["tile_streamed", "lds_staged"])
def test_primal_weight_integral_equals_model(engine, variant):
    """The proportional-integral primal-weight controller (primal_weight_integral > 0): 6 blocks
    cross several restarts, each of which moves the weight by theta e + ki sum(e); the device
    equals the model's run with the same gains."""
    for name, lp in lps():
        print(f"[case] step parity integral 0.1 {variant} {name}", flush=True)
        engine.upload(lp)
        r = engine.solve(vp.default_params(vp.MODE_RESTARTED, precision="fp64", tolerance=0.0, max_iterations=6 * K,
                                           check_frequency=K, primal_weight_smoothing=0.3,
                                           primal_weight_integral=0.1, **VARIANTS[variant]))
        x, _, y = engine.download()
        mx, my, restarts = model_point(lp, 6, smoothing=0.3, integral=0.1)
        assert r.restart_count == restarts and restarts >= 2, (r.restart_count, restarts)
        close(x, mx, (name, "x"))
        close(y, my, (name, "y"))


@pytest.mark.parametrize("blocks", [1, 3])
@pytest.mark.parametrize("variant", ["sliced_ell", "split_A4_At2_streamed", "tile_streamed"])
def test_reflection_half_equals_model(engine, variant, blocks):
    """Reflection 0.5: the x iterate then lives in its own vector (with reflection 1 the steps
    recover it from x_bar); both forms equal the model."""
    for name, lp in lps()[:2]:
        print(f"[case] step parity reflection 0.5 {variant} blocks={blocks} {name}", flush=True)
        engine.upload(lp)
        r = engine.solve(vp.default_params(vp.MODE_RESTARTED, precision="fp64", tolerance=0.0, reflection=0.5,
                                           max_iterations=blocks * K, check_frequency=K, **VARIANTS[variant]))
        x, _, y = engine.download()
        mx, my, restarts = model_point(lp, blocks, rho=0.5)
        assert r.restart_count == restarts
        close(x, mx, (name, "x"))
        close(y, my, (name, "y"))


@pytest.mark.parametrize("variant", sorted(VARIANTS))
def test_refine_every_check_equals_model(engine, variant):
    """refine_every_check = 1: the checks that do not restart move the FP64 centre to the check
    point and keep the Halpern epoch, whose anchor becomes a vector. 8 blocks cross restarts and
    such re-centrings; the device equals the model run with the same option."""
    for name, lp in lps():
        print(f"[case] step parity refine every check {variant} {name}", flush=True)
        engine.upload(lp)
        r = engine.solve(vp.default_params(vp.MODE_RESTARTED, precision="fp64", tolerance=0.0, max_iterations=8 * K,
                                           check_frequency=K, refine_every_check=1, **VARIANTS[variant]))
        x, _, y = engine.download()
        mx, my, restarts = model_point(lp, 8, refine_every_check=True)
        assert r.restart_count == restarts and 1 <= restarts < 8, (r.restart_count, restarts)
        close(x, mx, (name, "x"))
        close(y, my, (name, "y"))


@pytest.mark.parametrize("variant", ["tile_streamed", "split_A4_At2_streamed"])
def test_refine_every_check_reflection_half_equals_model(engine, variant):
    for name, lp in lps()[:2]:
        print(f"[case] step parity refine every check, reflection 0.5, {variant} {name}", flush=True)
        engine.upload(lp)
        r = engine.solve(vp.default_params(vp.MODE_RESTARTED, precision="fp64", tolerance=0.0, reflection=0.5,
                                           max_iterations=8 * K, check_frequency=K, refine_every_check=1,
                                           **VARIANTS[variant]))
        x, _, y = engine.download()
        mx, my, restarts = model_point(lp, 8, rho=0.5, refine_every_check=True)
        assert r.restart_count == restarts
        close(x, mx, (name, "x"))
        close(y, my, (name, "y"))


@pytest.mark.parametrize("lanes", [-1, 8, -5, -7])
def test_refine_every_check_batch_equals_model(engine, lanes):
    cases = lps()
    engine.upload_batch([lp for _, lp in cases])
    print(f"[case] batch step parity refine every check lanes={lanes}", flush=True)
    res = engine.solve_batch(vp.default_params(vp.MODE_RESTARTED, precision="fp64", tolerance=0.0,
                                               max_iterations=8 * K, check_frequency=K, lanes_A=lanes,
                                               lanes_At=lanes, refine_every_check=1))
    xs, ys = engine.download_batch()
    for (name, lp), r, x, y in zip(cases, res, xs, ys):
        mx, my, restarts = model_point(lp, 8, refine_every_check=True)
        assert r.restart_count == restarts
        close(x, mx, (name, "x"))
        close(y, my, (name, "y"))


@pytest.mark.parametrize("variant", ["tile_streamed", "split_A4_At2_streamed", "lds_staged_w16_256"])
def test_refine_every_check_short_blocks_equal_model(engine, variant):
    """Refinement at every check with blocks of 8 iterations (the centre moves every 8
    iterations, restarts come often): 24 blocks equal the model run with the same option."""
    k = 8
    for name, lp in lps()[:2]:
        print(f"[case] step parity refine every check K={k} {variant} {name}", flush=True)
        engine.upload(lp)
        r = engine.solve(vp.default_params(vp.MODE_RESTARTED, precision="fp64", tolerance=0.0, max_iterations=24 * k,
                                           check_frequency=k, refine_every_check=1, **VARIANTS[variant]))
        x, _, y = engine.download()
        m = M.solve(lp, K=k, tol=0.0, max_iters=24 * k, refine_every_check=True)
        assert m["iters"] == 24 * k and r.restart_count == m["restarts"], (r.restart_count, m["restarts"])
        close(x, m["x"], (name, "x"))
        close(y, m["y"], (name, "y"))
