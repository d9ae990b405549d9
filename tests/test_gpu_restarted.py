"""Restarted Halpern PDHG engine (VJ_MODE_RESTARTED): final-answer parity.

The bar (build-defined rel-KKT, FP64 on the original data, evaluated both by the device
audit and by the oracle on the downloaded point): rel-KKT <= 1e-6, and the objective within
1e-6 (relative) of the published / dual-simplex optimum."""
import dataclasses

import numpy as np
import pytest

import lp_cases
import oracle as O
import vajra_pdhg as vp

pytestmark = pytest.mark.gpu


def solve(engine, lp, tol=1e-6, max_iterations=200000, **kw):
    engine.upload(lp)
    p = vp.default_params(vp.MODE_RESTARTED, tolerance=tol, max_iterations=max_iterations, **kw)
    r = engine.solve(p)
    x, rd, y = engine.download()
    return r, x, rd, y


def check_point(lp, r, x, y, tol):
    k = O.rel_kkt(lp, x, y)
    # device audit agrees with the oracle's rel-KKT on the same point
    assert abs(k[0] - r.rel_kkt) <= 1e-9 + 1e-6 * k[0], (k, r.rel_kkt)
    assert k[0] <= tol, k
    return k


@pytest.mark.parametrize("lanes", [0, -3], ids=["autotuned", "sliced_ell"])
@pytest.mark.parametrize("name", sorted(lp_cases.kat_lps()))
def test_kat_fp64(engine, name, lanes):
    """The reference's known-answer LPs (tiny, with empty and singleton rows and fewer rows
    than one ELL slice), with the autotuned kernels and with the sliced-ELL kernels forced: rel-KKT
    1e-8 (oracle-checked) and the known optimum to 1e-6 relative."""
    lp, obj, xs = lp_cases.kat_lps()[name]
    print(f"[case] restarted fp64 kat {name} lanes={lanes}", flush=True)
    r, x, rd, y = solve(engine, lp, tol=1e-8, lanes_A=lanes, lanes_At=lanes)
    assert r.status == 1, r.as_dict()
    if lanes == -3:
        assert (r.lanes_A, r.lanes_At) == (-3, -3), r.as_dict()
    check_point(lp, r, x, y, 1e-8)
    got = float(np.dot(lp.c, x)) + lp.offset
    assert abs(got - obj) <= 1e-6 * (1 + abs(obj))


@pytest.mark.parametrize("prec", ["fp64", "fp32", "bf16", "fp16"])
def test_cfg0_to_1e6(engine, prec):
    lp = lp_cases.cfg0_lp()
    print(f"[case] restarted cfg0 {prec}", flush=True)
    r, x, rd, y = solve(engine, lp, precision=prec)
    assert r.status == 1, r.as_dict()
    k = check_point(lp, r, x, y, 1e-6)
    print(prec, r.iterations, r.restart_count, r.iteration_time, k[:4])


def test_fp32_refinement_to_1e8(engine):
    """FP32 iterates with FP64 refinement reach rel-KKT 1e-8 (oracle-checked) on the
    2,000 x 5,000 LP, and the objective agrees with the FP64 engine's at 1e-8 to 1e-7
    relative: the refinement restarts (on the FP64 rel-KKT of the checks) remove the error of
    the rounded FP32 data."""
    lp = lp_cases.cfg0_lp()
    print("[case] cfg0 fp32 to 1e-8", flush=True)
    r, x, rd, y = solve(engine, lp, tol=1e-8, precision="fp32")
    assert r.status == 1, r.as_dict()
    check_point(lp, r, x, y, 1e-8)
    r64, x64, _, _ = solve(engine, lp, tol=1e-8)
    assert r64.status == 1, r64.as_dict()
    o32, o64 = float(np.dot(lp.c, x)), float(np.dot(lp.c, x64))
    assert abs(o32 - o64) <= 1e-7 * (1 + abs(o64)), (o32, o64)


def test_restart_drift_parameter_range(engine):
    """restart_drift in (0, 1] would restart at every check: rejected as an invalid
    argument; 0 (default), negative (off) and factors above 1 are accepted."""
    engine.upload(lp_cases.cfg0_lp())
    for bad in (0.5, 1.0, float("nan")):
        with pytest.raises(vp.VjError):
            engine.set_params(vp.default_params(vp.MODE_RESTARTED, restart_drift=bad))
    for ok in (0.0, -1.0, 1.5, 4.0):
        engine.set_params(vp.default_params(vp.MODE_RESTARTED, restart_drift=ok))
    for bad in (-2, 2):
        with pytest.raises(vp.VjError):
            engine.set_params(vp.default_params(vp.MODE_RESTARTED, refine_every_check=bad))
    for ok in (-1, 0, 1):
        engine.set_params(vp.default_params(vp.MODE_RESTARTED, refine_every_check=ok))
    for bad in (512, 7, -1):
        with pytest.raises(vp.VjError):
            engine.set_params(vp.default_params(vp.MODE_RESTARTED, stage_threads=bad))
    for ok in (0, 256, 1024):
        engine.set_params(vp.default_params(vp.MODE_RESTARTED, stage_threads=ok))


def test_cfg0_objective_against_highs(engine):
    """The FP64 engine's objective on the 2,000 x 5,000 LP against the optimum SciPy-HiGHS
    found for the same LP (the reference's own dual simplex fails on it: status 9, recorded
    in the golden metadata). At rel-KKT 1e-8 the objective agrees to 1e-6 relative."""
    import json
    import os
    meta = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "golden_meta.json")))
    opt = meta["cfg0_highs_objective"]
    lp = lp_cases.cfg0_lp()
    print("[case] cfg0 fp64 objective vs HiGHS", flush=True)
    r, x, rd, y = solve(engine, lp, tol=1e-8)
    assert r.status == 1, r.as_dict()
    check_point(lp, r, x, y, 1e-8)
    got = float(np.dot(lp.c, x)) + lp.offset
    assert abs(got - opt) <= 1e-6 * (1 + abs(opt)), (got, opt)


@pytest.mark.parametrize("prec,lanes", [("fp64", 0), ("fp64", -3), ("fp32", -4)])
@pytest.mark.parametrize("name", ["afiro", "crude_blend"])
def test_mps_fixtures(engine, name, prec, lanes):
    """Netlib afiro and crude_blend to rel-KKT 1e-8 against their published optima, with the
    autotuned kernels and with the sliced-ELL kernels forced (FP64, and FP32 records with FP64
    refinement)."""
    lp = lp_cases.mps_lp(name)
    opt = lp_cases.MPS_OPTIMA[name]
    print(f"[case] mps {name} {prec} lanes={lanes}", flush=True)
    r, x, rd, y = solve(engine, lp, tol=1e-8, precision=prec, lanes_A=lanes, lanes_At=lanes)
    assert r.status == 1, r.as_dict()
    if lanes < 0:
        assert (r.lanes_A, r.lanes_At) == (lanes, lanes), r.as_dict()
    check_point(lp, r, x, y, 1e-8)
    got = float(np.dot(lp.c, x)) + lp.offset
    assert abs(got - opt) <= 1e-6 * (1 + abs(opt))


def test_graph_and_stream_paths_identical(engine):
    """The graph-replayed block and the directly launched block run the same kernels in the
    same order: bit-identical answers; a rerun is bit-identical too (no float atomics)."""
    lp = lp_cases.cfg0_lp()
    r1, x1, _, y1 = solve(engine, lp, use_graph=1)
    r2, x2, _, y2 = solve(engine, lp, use_graph=0)
    r3, x3, _, y3 = solve(engine, lp, use_graph=1)
    assert r1.iterations == r2.iterations == r3.iterations
    assert np.array_equal(x1, x2) and np.array_equal(y1, y2)
    assert np.array_equal(x1, x3) and np.array_equal(y1, y3)


def test_batch_matches_single(engine):
    lps = [vp.gen_random_lp(1000 + i, m, int(2.5 * m)) for i, m in enumerate([500, 800, 1300, 650])]
    engine.upload_batch(lps)
    res = engine.solve_batch(vp.default_params(vp.MODE_RESTARTED, tolerance=1e-6, max_iterations=200000))
    xs, ys = engine.download_batch()
    for lp, r, x, y in zip(lps, res, xs, ys):
        assert r.status == 1, r.as_dict()
        k = O.rel_kkt(lp, x, y)
        assert k[0] <= 1e-6
        assert abs(k[0] - r.rel_kkt) <= 1e-9 + 1e-6 * k[0]


@pytest.mark.parametrize("prec,sa,sat", [("fp64", 8, 4), ("fp32", 4, 2), ("bf16", 2, 8), ("fp16", 4, 2)])
def test_column_split_spmv(engine, prec, sa, sat):
    """The XCD column-split SpMV (block partial sums + ordered finish) solves to the same
    tolerance as the fused kernels, deterministically (two runs bit-identical)."""
    lp = lp_cases.cfg0_lp()
    print(f"[case] split {prec} SA={sa} SAt={sat}", flush=True)
    r1, x1, _, y1 = solve(engine, lp, precision=prec, col_splits_A=sa, col_splits_At=sat)
    assert (r1.col_splits_A, r1.col_splits_At) == (sa, sat)
    assert r1.status == 1, r1.as_dict()
    check_point(lp, r1, x1, y1, 1e-6)
    r2, x2, _, y2 = solve(engine, lp, precision=prec, col_splits_A=sa, col_splits_At=sat)
    assert r1.iterations == r2.iterations and np.array_equal(x1, x2) and np.array_equal(y1, y2)


def long_row_lp(seed=7, m=10000, n=24000, k=4):
    """Feasible, bounded LP (the random construction: b = A x0, c = A^T y0 + s0) whose row 0
    is dense over all n columns and whose column 0 is dense over all m rows, so that tiles of
    one row longer than the 4096-entry LDS capacity occur in A and A^T, also per column block
    of the split layouts (n/4 = 6000 and m/2 = 5000 entries)."""
    rng = np.random.default_rng(seed)
    cols = [np.union1d(rng.choice(m, k, replace=False), [0]) for _ in range(n)]
    cols[0] = np.arange(m)
    col_ptr = np.zeros(n + 1, np.int64)
    col_ptr[1:] = np.cumsum([len(c) for c in cols])
    row_idx = np.concatenate(cols).astype(np.int32)
    vals = rng.standard_normal(len(row_idx))
    lp = vp.LP(m, n, col_ptr, row_idx, vals, np.zeros(n), np.zeros(n), np.full(n, 10.0), np.zeros(m), np.zeros(m))
    x0 = rng.uniform(0, 1, n)
    y0 = rng.standard_normal(m)
    A = (lp.col_ptr, lp.row_idx, lp.vals)
    b = np.zeros(m)
    c = rng.uniform(0, 1, n)
    for j in range(n):
        s, e = A[0][j], A[0][j + 1]
        b[A[1][s:e]] += A[2][s:e] * x0[j]
        c[j] += np.dot(A[2][s:e], y0[A[1][s:e]])
    lp.row_lower[:] = b
    lp.row_upper[:] = b
    lp.c[:] = c
    return lp


_LONG = {}


@pytest.mark.parametrize("prec,lanes,sa,sat", [
    ("fp32", -2, 1, 1), ("fp32", -1, 1, 1), ("fp32", -2, 4, 2), ("fp32", -1, 4, 2),
    ("fp64", -1, 4, 2), ("fp64", -1, 1, 1), ("bf16", -2, 2, 2)])
def test_tile_stream_variants(engine, prec, lanes, sa, sat):
    """Each tile-streamed kernel (fused, column-split, with separate index/value arrays or
    with FP32 (index, value) records), forced on an LP with rows and columns longer than the
    LDS tile: solves to rel-KKT 1e-6 (oracle-checked), the objective agrees with the
    autotuned FP64 solve, and a rerun is bit-identical."""
    lp = _LONG.get("lp")
    if lp is None:
        lp = _LONG["lp"] = long_row_lp()
    print(f"[case] tile-stream {prec} lanes={lanes} SA={sa} SAt={sat}", flush=True)
    kw = dict(precision=prec, lanes_A=lanes, lanes_At=lanes, col_splits_A=sa, col_splits_At=sat)
    r1, x1, _, y1 = solve(engine, lp, **kw)
    assert (r1.col_splits_A, r1.col_splits_At) == (sa, sat)
    assert r1.status == 1, r1.as_dict()
    check_point(lp, r1, x1, y1, 1e-6)
    r2, x2, _, y2 = solve(engine, lp, **kw)
    assert r1.iterations == r2.iterations and np.array_equal(x1, x2) and np.array_equal(y1, y2)
    if "obj" not in _LONG:
        r0, x0, _, _ = solve(engine, lp, tol=1e-8)
        assert r0.status == 1
        _LONG["obj"] = float(np.dot(lp.c, x0))
    o0, o1 = _LONG["obj"], float(np.dot(lp.c, x1))
    assert abs(o1 - o0) <= 2e-5 * (1 + abs(o0)), (o1, o0)


@pytest.mark.parametrize("prec,lanes", [("fp32", -4), ("fp32", -3), ("fp64", -3), ("bf16", -4), ("bf16", -3),
                                        ("fp16", -3)])
def test_sliced_ell_variants(engine, prec, lanes):
    """The sliced-ELL step kernels (slices of 64 rows, one thread per row; -4 reads FP32
    (index, value) records) forced on the 2,000 x 5,000 LP, whose rows of A vary in length so
    that slices carry padding: solves to rel-KKT 1e-6 (oracle-checked), reports the ELL kernel
    (records only in FP32), the objective agrees with the tile-streamed solve of the same
    precision to 2e-5, and a rerun is bit-identical."""
    lp = lp_cases.cfg0_lp()
    print(f"[case] sliced ELL {prec} lanes={lanes}", flush=True)
    kw = dict(precision=prec, lanes_A=lanes, lanes_At=lanes, col_splits_A=1, col_splits_At=1)
    r1, x1, _, y1 = solve(engine, lp, **kw)
    want = -4 if lanes == -4 and prec == "fp32" else -3
    assert (r1.lanes_A, r1.lanes_At) == (want, want), r1.as_dict()
    assert r1.status == 1, r1.as_dict()
    check_point(lp, r1, x1, y1, 1e-6)
    r2, x2, _, y2 = solve(engine, lp, **kw)
    assert r1.iterations == r2.iterations and np.array_equal(x1, x2) and np.array_equal(y1, y2)
    r0, x0, _, _ = solve(engine, lp, precision=prec, lanes_A=-1, lanes_At=-1, col_splits_A=1, col_splits_At=1)
    assert r0.status == 1 and (r0.lanes_A, r0.lanes_At) == (0, 0)
    o0, o1 = float(np.dot(lp.c, x0)), float(np.dot(lp.c, x1))
    assert abs(o1 - o0) <= 2e-5 * (1 + abs(o0)), (o1, o0)


STRUCTURED = [(0, 12, 30), (1, 30, 12), (2, 1, 9), (3, 9, 1), (4, 40, 40), (5, 60, 150), (6, 150, 60), (7, 3, 3)]


@pytest.mark.parametrize("prec,lanes", [("fp64", 0), ("fp32", 0), ("fp64", -3)])
@pytest.mark.parametrize("seed,m,n", STRUCTURED)
def test_structured_lps_against_highs(engine, seed, m, n, prec, lanes):
    """Feasible, bounded LPs with empty rows and columns, fixed, free and one-sided columns,
    equality, ranged and one-sided rows, tiny coefficients, offsets and both senses: the
    restarted engine reaches rel-KKT 1e-8 (oracle-checked on the downloaded point) and the
    objective of SciPy's HiGHS to 1e-6 relative."""
    lp = lp_cases.structured_feasible_lp(seed, m, n)
    opt = lp_cases.highs_optimum(lp)
    assert opt is not None
    print(f"[case] structured seed={seed} {m}x{n} {prec} lanes={lanes}", flush=True)
    r, x, rd, y = solve(engine, lp, tol=1e-8, precision=prec, lanes_A=lanes, lanes_At=lanes)
    assert r.status == 1, r.as_dict()
    check_point(lp, r, x, y, 1e-8)
    got = float(np.dot(lp.c, x)) + lp.offset
    assert abs(got - opt) <= 1e-6 * (1 + abs(opt)), (got, opt)


@pytest.mark.parametrize("prec,lanes", [("fp64", 0), ("fp32", 0), ("fp64", -5), ("fp32", -6), ("fp32", -7)])
def test_structured_batch_against_highs(engine, prec, lanes):
    """The eight structured LPs (1x9 and 9x1 among them) solved as one block-diagonal batch,
    twice over in mixed order, so that tiles span several instances of different shapes:
    every instance reaches rel-KKT 1e-8 (oracle-checked) and its HiGHS objective to 1e-6.
    Also with the LDS-staged kernels forced (their tiles hold one instance each)."""
    lps = [lp_cases.structured_feasible_lp(*s) for s in STRUCTURED]
    lps = lps + lps[::-1]
    opts = [lp_cases.highs_optimum(lp) for lp in lps]
    print(f"[case] structured batch {prec} lanes={lanes} ({len(lps)} instances)", flush=True)
    engine.upload_batch(lps)
    res = engine.solve_batch(vp.default_params(vp.MODE_RESTARTED, precision=prec, tolerance=1e-8,
                                               max_iterations=200000, lanes_A=lanes, lanes_At=lanes))
    xs, ys = engine.download_batch()
    if lanes:
        assert all((r.lanes_A, r.lanes_At) == (lanes, lanes) for r in res), res[0].as_dict()
    for lp, opt, r, x, y in zip(lps, opts, res, xs, ys):
        assert r.status == 1, r.as_dict()
        k = O.rel_kkt(lp, x, y)
        assert k[0] <= 1e-8 and abs(k[0] - r.rel_kkt) <= 1e-11 + 1e-6 * k[0], (k, r.rel_kkt)
        got = float(np.dot(lp.c, x)) + lp.offset
        assert abs(got - opt) <= 1e-6 * (1 + abs(opt)), (got, opt)


@pytest.mark.parametrize("threads", [1024, 256])
@pytest.mark.parametrize("prec,stream,staged", [("fp64", -1, -5), ("fp32", -2, -6), ("bf16", -1, -5),
                                                ("fp64", -1, -7), ("fp32", -2, -7), ("bf16", -1, -7),
                                                ("fp16", -1, -5), ("fp16", -1, -7)])
def test_staged_equals_tile_streamed(engine, prec, stream, staged, threads):
    """The LDS-staged steps form the same products and add each row in index order, as the
    tile-streamed steps do, so one check block of a batch (before the first restart decision,
    whose residual partials the two kernels sum in different tile orders) returns the same bits.
    The batch mixes random and structured LPs, with windows wider than one LDS chunk (the
    7,000 x 17,500 instance copies its x_bar window in two chunks in FP32, three in FP64).
    Lanes -7 reads 16-bit column offsets from each tile's window start (with BF16 and FP16
    values, packed with the value into one 4-byte record). Both workgroup shapes (1,024 and
    256 threads, with tiles and LDS windows of their own sizes) give the same bits."""
    lps = [vp.gen_random_lp(4100 + i, m, int(2.5 * m)) for i, m in enumerate([500, 7000, 1300])]
    lps += [lp_cases.structured_feasible_lp(*s) for s in STRUCTURED[:3]]
    engine.upload_batch(lps)
    out = []
    for lanes in (stream, staged):
        print(f"[case] staged vs tile-streamed {prec} lanes={lanes} threads={threads}", flush=True)
        res = engine.solve_batch(vp.default_params(vp.MODE_RESTARTED, precision=prec, tolerance=0.0,
                                                   max_iterations=64, check_frequency=64, lanes_A=lanes,
                                                   lanes_At=lanes, stage_threads=threads))
        want = 0 if lanes == stream else staged   # the result reports the tile-streamed kernel as 0
        assert all((r.lanes_A, r.lanes_At) == (want, want) and r.iterations == 64 for r in res), res[0].as_dict()
        out.append(engine.download_batch())
    (xs0, ys0), (xs1, ys1) = out
    for x0, y0, x1, y1 in zip(xs0, ys0, xs1, ys1):
        assert np.array_equal(x0, x1) and np.array_equal(y0, y1)


@pytest.mark.parametrize("prec,lanes,threads", [("fp32", -7, 1024), ("fp64", -5, 1024), ("fp16", -7, 256),
                                                ("fp64", -5, 256)])
def test_staged_batch_graph_and_stream_identical(engine, prec, lanes, threads):
    """A batch whose instances stop at very different iteration counts, with the LDS-staged
    kernels: as instances finish, the host drops their tiles from the staged launches (and
    captures the block graph again with the smaller grids). Graph replays and direct launches
    prune at the same blocks, so both return the same bits, and every instance is checked by
    the oracle at its own tolerance."""
    lps = [vp.gen_random_lp(4200 + i, m, int(2.5 * m)) for i, m in enumerate([300, 2500, 600, 4000, 1200, 5000])]
    lps += [lp_cases.structured_feasible_lp(*s) for s in STRUCTURED[:4]]
    engine.upload_batch(lps)
    out = []
    for graph in (1, 0):
        print(f"[case] staged batch {prec} lanes={lanes} threads={threads} use_graph={graph}", flush=True)
        res = engine.solve_batch(vp.default_params(vp.MODE_RESTARTED, precision=prec, tolerance=1e-7,
                                                   max_iterations=200000, lanes_A=lanes, lanes_At=lanes,
                                                   use_graph=graph, stage_threads=threads))
        xs, ys = engine.download_batch()
        for lp, r, x, y in zip(lps, res, xs, ys):
            assert r.status == 1 and (r.lanes_A, r.lanes_At) == (lanes, lanes), r.as_dict()
            k = O.rel_kkt(lp, x, y)
            assert k[0] <= 1e-7 and abs(k[0] - r.rel_kkt) <= 1e-10 + 1e-6 * k[0], (k, r.rel_kkt)
        out.append(([r.iterations for r in res], xs, ys))
    (it0, xs0, ys0), (it1, xs1, ys1) = out
    print(f"[case] iterations per instance {it0}", flush=True)
    assert it0 == it1 and max(it0) >= 4 * min(it0)
    for x0, y0, x1, y1 in zip(xs0, ys0, xs1, ys1):
        assert np.array_equal(x0, x1) and np.array_equal(y0, y1)


def test_fp16_range_checked(engine):
    """FP16 matrix values: column 0 of the 2,000 x 5,000 LP scaled by 1e5 (cost scaled with it,
    bounds divided: the same LP in other units) holds entries beyond the FP16 range (65504)
    once scaling is off, and the solve is refused with an error rather than run with infinite
    entries. The context stays usable: the unscaled LP then solves in FP16 to rel-KKT 1e-6
    (oracle-checked)."""
    lp = lp_cases.cfg0_lp()
    s = 1e5
    vals = lp.vals.copy()
    vals[lp.col_ptr[0]:lp.col_ptr[1]] *= s
    c, lc, uc = lp.c.copy(), lp.col_lower.copy(), lp.col_upper.copy()
    c[0] *= s
    lc[0] /= s
    uc[0] /= s
    lp2 = dataclasses.replace(lp, vals=vals, c=c, col_lower=lc, col_upper=uc, _csr=None)
    assert np.max(np.abs(vals)) > 65504
    print("[case] fp16 range check", flush=True)
    engine.upload(lp2)
    with pytest.raises(vp.VjError) as e:
        engine.solve(vp.default_params(vp.MODE_RESTARTED, precision="fp16", tolerance=1e-6, enable_scaling=0))
    assert "FP16" in str(e.value)
    r, x, rd, y = solve(engine, lp, precision="fp16")
    assert r.status == 1, r.as_dict()
    check_point(lp, r, x, y, 1e-6)


def test_infeasible_lps_never_reported_optimal(engine):
    """LPs that SciPy-HiGHS finds infeasible: the restarted engine (which, like the
    reference's PDHG, has no infeasibility certificate) stops at its iteration limit and
    never reports them optimal, singly or inside a batch, where the feasible instances
    still reach rel-KKT 1e-8 and their HiGHS objectives to 1e-6."""
    bad = [lp_cases.structured_lp(s, m, n) for s, m, n in [(1, 30, 12), (4, 40, 40)]]
    for lp in bad:
        assert lp_cases.highs_optimum(lp) is None
        print(f"[case] infeasible {lp.m}x{lp.n}", flush=True)
        r, x, rd, y = solve(engine, lp, max_iterations=20000)
        assert r.status != 1, r.as_dict()
    good = [lp_cases.structured_feasible_lp(*s) for s in STRUCTURED[:4]]
    lps = [good[0], bad[0], good[1], good[2], bad[1], good[3]]
    print("[case] batch with infeasible instances", flush=True)
    engine.upload_batch(lps)
    res = engine.solve_batch(vp.default_params(vp.MODE_RESTARTED, tolerance=1e-8, max_iterations=100000))
    xs, ys = engine.download_batch()
    for lp, r, x, y in zip(lps, res, xs, ys):
        opt = lp_cases.highs_optimum(lp)
        if opt is None:
            assert r.status != 1, r.as_dict()
            continue
        assert r.status == 1, r.as_dict()
        assert O.rel_kkt(lp, x, y)[0] <= 1e-8
        got = float(np.dot(lp.c, x)) + lp.offset
        assert abs(got - opt) <= 1e-6 * (1 + abs(opt)), (got, opt)


@pytest.mark.parametrize("where", ["c", "vals"])
def test_nan_input_is_a_numerical_error(engine, where):
    """A NaN in the costs or the matrix ends the solve with NUMERICAL_ERROR (the reference's
    kNumericalError for non-finite iterates, pdhg.cpp:215-219), not with a point."""
    lp = lp_cases.structured_feasible_lp(0, 12, 30)
    if where == "c":
        lp.c[4] = np.nan
    else:
        lp.vals[3] = np.nan
    print(f"[case] NaN in {where}", flush=True)
    engine.upload(lp)
    r = engine.solve(vp.default_params(vp.MODE_RESTARTED, tolerance=1e-6, max_iterations=5000))
    assert r.status == 9, r.as_dict()


@pytest.mark.parametrize("decades", [1.0, 3.0])
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("seed,m,n", [(0, 12, 30), (1, 30, 12), (4, 40, 40), (5, 60, 150), (6, 150, 60)])
def test_badly_scaled_lps(engine, seed, m, n, prec, decades):
    """The feasible structured LPs with rows and columns rescaled by 10^U[-d,d], d = 1 and 3 (an exact
    equivalence, so the optimum is unchanged): Ruiz and Pock-Chambolle scaling on the device
    absorb it, and the restarted engine still reaches rel-KKT 1e-8 (oracle-checked on the
    rescaled LP) and the HiGHS objective of the original LP to 1e-6. These LPs hold a row and
    a column of tiny coefficients only; without a bound on the equilibration factors their
    bounds / cost were inflated by up to 1e10 and the bound/objective rescaling stalled the
    iteration (rel-KKT frozen near 1)."""
    lp0 = lp_cases.structured_feasible_lp(seed, m, n)
    opt = lp_cases.highs_optimum(lp0)
    lp = lp_cases.rescaled_lp(lp0, seed, decades)
    print(f"[case] rescaled seed={seed} {m}x{n} {prec} 10^+-{decades}", flush=True)
    r, x, rd, y = solve(engine, lp, tol=1e-8, precision=prec)
    assert r.status == 1, r.as_dict()
    check_point(lp, r, x, y, 1e-8)
    got = float(np.dot(lp.c, x)) + lp.offset
    assert abs(got - opt) <= 1e-6 * (1 + abs(opt)), (got, opt)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("seed,m,n,decades", [(11, 2, 64, 0.0), (11, 2, 64, 2.0), (33, 15, 6, 0.0)])
def test_near_empty_columns_with_costs(engine, seed, m, n, decades, prec):
    """LPs with a column whose only coefficients are tiny (1e-11 to 1e-8) but whose cost is
    not: equilibration scales such a column to its bound (1e6), and its inflated cost used to
    dominate the objective norm (crushing every other cost in the rescaling) and the initial
    primal weight, so the rows never became active (duals stuck at 0). The rescaling and the
    primal weight now leave such rows and columns out; the engine reaches rel-KKT 1e-8 and
    the HiGHS objective to 1e-6."""
    lp0 = lp_cases.structured_feasible_lp(seed, m, n)
    lp = lp_cases.rescaled_lp(lp0, seed, decades) if decades else lp0
    opt = lp_cases.highs_optimum(lp)
    assert opt is not None
    print(f"[case] near-empty columns seed={seed} {m}x{n} 10^+-{decades} {prec}", flush=True)
    r, x, rd, y = solve(engine, lp, tol=1e-8, precision=prec)
    assert r.status == 1, r.as_dict()
    check_point(lp, r, x, y, 1e-8)
    got = float(np.dot(lp.c, x)) + lp.offset
    assert abs(got - opt) <= 1e-6 * (1 + abs(opt)), (got, opt)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_badly_scaled_batch(engine, prec):
    """The rescaled LPs (10^+-1 and 10^+-3, with their near-empty rows and columns) and the
    near-empty-column LPs as one batch: per-instance equilibration bounds and per-instance
    rescaling norms; every instance reaches rel-KKT 1e-8 and its HiGHS objective to 1e-6."""
    lps = []
    for dec in (1.0, 3.0):
        for s, m, n in [(0, 12, 30), (1, 30, 12), (4, 40, 40), (5, 60, 150), (6, 150, 60)]:
            lps.append(lp_cases.rescaled_lp(lp_cases.structured_feasible_lp(s, m, n), s, dec))
    lps.append(lp_cases.structured_feasible_lp(11, 2, 64))
    lps.append(lp_cases.rescaled_lp(lp_cases.structured_feasible_lp(11, 2, 64), 11, 2.0))
    lps.append(lp_cases.structured_feasible_lp(33, 15, 6))
    opts = [lp_cases.highs_optimum(lp) for lp in lps]
    print(f"[case] badly scaled batch {prec} ({len(lps)} instances)", flush=True)
    engine.upload_batch(lps)
    res = engine.solve_batch(vp.default_params(vp.MODE_RESTARTED, precision=prec, tolerance=1e-8,
                                               max_iterations=200000))
    xs, ys = engine.download_batch()
    for lp, opt, r, x, y in zip(lps, opts, res, xs, ys):
        assert r.status == 1, r.as_dict()
        k = O.rel_kkt(lp, x, y)
        assert k[0] <= 1e-8 and abs(k[0] - r.rel_kkt) <= 1e-11 + 1e-6 * k[0], (k, r.rel_kkt)
        got = float(np.dot(lp.c, x)) + lp.offset
        assert abs(got - opt) <= 1e-6 * (1 + abs(opt)), (got, opt)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_default_solves_reproducible(engine, prec):
    """Repeated solves with the default (autotuned) kernel choices agree bit for bit, as the
    reference asserts for its repeated solves (tests/test_reliability.cpp:343-372): on an LP
    above the autotuner's 200K-nonzero threshold the first solve times the candidates, later
    solves of the same upload reuse that decision, and every kernel sums in a fixed order."""
    lp = vp.gen_random_lp(0x5EED2026 + 11, 50_000, 125_000, 6, 0.1, 3.0)
    assert lp.nnz > 200_000
    engine.upload(lp)
    runs = []
    for rep in range(3):
        print(f"[case] reproducible default solve {prec} #{rep}", flush=True)
        r = engine.solve(vp.default_params(vp.MODE_RESTARTED, precision=prec, tolerance=1e-6, max_iterations=400000))
        x, _, y = engine.download()
        assert r.status == 1, r.as_dict()
        runs.append((r.iterations, r.lanes_A, r.lanes_At, r.col_splits_A, r.col_splits_At, r.check_interval, x, y))
    for other in runs[1:]:
        assert other[:6] == runs[0][:6], (other[:6], runs[0][:6])
        assert np.array_equal(other[6], runs[0][6]) and np.array_equal(other[7], runs[0][7])
    assert O.rel_kkt(lp, runs[0][6], runs[0][7])[0] <= 1e-6


def _hard_small_lps():
    lps = [lp_cases.structured_feasible_lp(*s) for s in STRUCTURED]
    for s, m, n in [(0, 12, 30), (5, 60, 150), (6, 150, 60)]:
        lps.append(lp_cases.rescaled_lp(lp_cases.structured_feasible_lp(s, m, n), s, 3.0))
    lps.append(lp_cases.rescaled_lp(lp_cases.structured_feasible_lp(11, 2, 64), 11, 2.0))
    return lps


@pytest.mark.parametrize("prec", ["fp64", "fp32", "fp16", "bf16"])
def test_refine_every_check_hard_batch(engine, prec):
    """Refinement at every check (refine_every_check = 1) on the structured LPs (empty, fixed,
    free and one-sided rows and columns, offsets, both senses), rows and columns rescaled by
    10^+-3 and near-empty columns with costs, as one batch: every instance reaches rel-KKT 1e-8
    (oracle-checked on the downloaded point) and its HiGHS objective to 1e-6, in every value
    precision."""
    lps = _hard_small_lps()
    opts = [lp_cases.highs_optimum(lp) for lp in lps]
    print(f"[case] refine every check, hard batch {prec} ({len(lps)} instances)", flush=True)
    engine.upload_batch(lps)
    res = engine.solve_batch(vp.default_params(vp.MODE_RESTARTED, precision=prec, tolerance=1e-8,
                                               max_iterations=400000, refine_every_check=1))
    xs, ys = engine.download_batch()
    print(f"[case] max iterations {max(r.iterations for r in res)}", flush=True)
    for lp, opt, r, x, y in zip(lps, opts, res, xs, ys):
        assert r.status == 1, r.as_dict()
        k = O.rel_kkt(lp, x, y)
        assert k[0] <= 1e-8 and abs(k[0] - r.rel_kkt) <= 1e-11 + 1e-6 * k[0], (k, r.rel_kkt)
        got = float(np.dot(lp.c, x)) + lp.offset
        assert abs(got - opt) <= 1e-6 * (1 + abs(opt)), (got, opt)


def test_refine_every_check_failure_statuses(engine):
    """With refinement at every check an infeasible LP still ends at its iteration limit, never
    OPTIMAL, and a NaN in the costs still ends the solve with NUMERICAL_ERROR."""
    bad = lp_cases.structured_lp(1, 30, 12)
    assert lp_cases.highs_optimum(bad) is None
    print("[case] refine every check: infeasible fp32", flush=True)
    r, *_ = solve(engine, bad, max_iterations=20000, precision="fp32", refine_every_check=1)
    assert r.status != 1, r.as_dict()
    lp = lp_cases.structured_feasible_lp(0, 12, 30)
    lp.c[4] = np.nan
    print("[case] refine every check: NaN cost fp32", flush=True)
    r, *_ = solve(engine, lp, max_iterations=5000, precision="fp32", refine_every_check=1)
    assert r.status == 9, r.as_dict()


@pytest.mark.parametrize("threads", [1024, 256])
def test_batch_benchmark_kernels_names_the_staged_shape(engine, threads):
    """vj_benchmark_kernels after a batch solve times the staged steps the solve ran and names
    them with their workgroup size; the timing leaves the solved state in place (a second
    download returns the same point)."""
    lps = [vp.gen_random_lp(4300 + i, m, int(2.5 * m)) for i, m in enumerate([400, 900])]
    engine.upload_batch(lps)
    print(f"[case] batch benchmark_kernels staged threads={threads}", flush=True)
    engine.solve_batch(vp.default_params(vp.MODE_RESTARTED, precision="fp32", tolerance=1e-6, max_iterations=20000,
                                         lanes_A=-7, lanes_At=-7, stage_threads=threads))
    xs0, ys0 = engine.download_batch()
    ks = engine.benchmark_kernels(reps=3)
    names = [k["name"] for k in ks]
    assert len(ks) == 2 and all(n.startswith("staged_") and n.endswith(f",{threads}>") for n in names), names
    assert all(k["avg_us"] > 0 and k["bytes_per_launch"] > 0 for k in ks), ks
    xs1, ys1 = engine.download_batch()
    for a, b in zip(xs0 + ys0, xs1 + ys1):
        assert np.array_equal(a, b)


def test_bf16_values_to_1e8_at_100k(engine):
    """BF16 matrix values (FP32 iterates, FP64 refinement at every check, drift restarts of
    factor 4: the defaults for 16-bit values) to rel-KKT 1e-8 on a 100,000 x 200,000 LP of the
    benchmark family (10% of the columns scaled by 10^U[-3,3]): oracle-checked, and the
    objective within 1e-7 of the FP64 engine's at the same tolerance."""
    lp = vp.gen_random_lp(0x5EED2026 + 11, 100000, 200000, 6, 0.1, 3.0)
    print("[case] bf16 values to 1e-8, 100000 x 200000", flush=True)
    r, x, _, y = solve(engine, lp, tol=1e-8, max_iterations=3000000, precision="bf16", time_limit=150.0)
    print(f"[case] bf16: iterations {r.iterations}, restarts {r.restart_count}, K {r.check_interval}", flush=True)
    assert r.status == 1, r.as_dict()
    check_point(lp, r, x, y, 1e-8)
    r64, x64, _, _ = solve(engine, lp, tol=1e-8, max_iterations=3000000, precision="fp64", time_limit=150.0)
    assert r64.status == 1, r64.as_dict()
    o16, o64 = float(np.dot(lp.c, x)) + lp.offset, float(np.dot(lp.c, x64)) + lp.offset
    assert abs(o16 - o64) <= 1e-7 * (1 + abs(o64)), (o16, o64)


def _long_row_lp(seed, m=30, n=5000):
    """A feasible, bounded LP whose row 0 holds every column (5,000 entries: longer than a
    256-thread staged tile, 4,096 FP32 / 2,048 FP64 entries, and than a tile-streamed tile,
    4,096), the other rows about 500 entries each."""
    rng = np.random.default_rng(seed)
    A = np.zeros((m, n))
    for j in range(n):
        rows = rng.choice(np.arange(1, m), 3, replace=False)
        A[rows, j] = rng.standard_normal(3)
    A[0, :] = rng.standard_normal(n)
    b = A @ rng.uniform(0.0, 1.0, n)
    c = A.T @ rng.standard_normal(m) + rng.uniform(0.0, 1.0, n)
    return vp.LP.from_dense(A, c, np.zeros(n), np.full(n, 10.0), b, b)


@pytest.mark.parametrize("prec,stream,staged", [("fp64", -1, -5), ("fp32", -2, -7)])
def test_staged_long_row_equals_tile_streamed(engine, prec, stream, staged):
    """A row longer than a tile is summed by the whole workgroup (a strided pass and a tree):
    in 256-thread staged workgroups the same order as the tile-streamed kernels' 256-thread
    long-row sum, so one check block of a batch holding such a row gives the same bits."""
    lps = [_long_row_lp(5), vp.gen_random_lp(4400, 800, 2000)]
    engine.upload_batch(lps)
    out = []
    for lanes in (stream, staged):
        print(f"[case] long row {prec} lanes={lanes}", flush=True)
        res = engine.solve_batch(vp.default_params(vp.MODE_RESTARTED, precision=prec, tolerance=0.0,
                                                   max_iterations=64, check_frequency=64, lanes_A=lanes,
                                                   lanes_At=lanes, stage_threads=256))
        assert all(r.iterations == 64 for r in res), res[0].as_dict()
        out.append(engine.download_batch())
    (xs0, ys0), (xs1, ys1) = out
    for x0, y0, x1, y1 in zip(xs0, ys0, xs1, ys1):
        assert np.array_equal(x0, x1) and np.array_equal(y0, y1)


@pytest.mark.parametrize("threads", [1024, 256])
def test_staged_long_row_batch_against_highs(engine, threads):
    """The batch with the long row solved with the staged kernels in either workgroup shape
    (1,024 threads: the row fits a tile and one lane adds it; 256: the workgroup adds it):
    every instance reaches rel-KKT 1e-8 (oracle) and the long-row LP its HiGHS objective."""
    lps = [_long_row_lp(6), vp.gen_random_lp(4401, 800, 2000)]
    opt = lp_cases.highs_optimum(lps[0])
    assert opt is not None
    engine.upload_batch(lps)
    print(f"[case] long-row batch staged threads={threads}", flush=True)
    res = engine.solve_batch(vp.default_params(vp.MODE_RESTARTED, precision="fp64", tolerance=1e-8,
                                               max_iterations=200000, lanes_A=-5, lanes_At=-5,
                                               stage_threads=threads))
    xs, ys = engine.download_batch()
    for lp, r, x, y in zip(lps, res, xs, ys):
        assert r.status == 1, r.as_dict()
        assert O.rel_kkt(lp, x, y)[0] <= 1e-8
    got = float(np.dot(lps[0].c, xs[0])) + lps[0].offset
    assert abs(got - opt) <= 1e-6 * (1 + abs(opt)), (got, opt)
