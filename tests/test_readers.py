"""MPS and LP readers of the C++ API (vajra-opt_amd/host/mps_reader.cpp, lp_reader.cpp)
against the reference's own readers compiled from its sources (oracle/_ref): the same file
must give the same Model, array for array and bit for bit. Inputs: every MPS fixture (the
Netlib set the reference ships with published optima, afiro, crude_blend), crude_blend.lp,
and files written here that exercise the formats' corner cases.

Also (GPU): the restarted FP64 engine reaches the published optimum of every Netlib
fixture to 1e-6 relative, at a relative KKT error of 1e-9 recomputed by the oracle."""
import glob
import json
import os
import subprocess

import numpy as np
import pytest

import lp_cases
import oracle as O
import vajra_pdhg as vp

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "vajra-opt_amd", "_build", "test_indus_api")
GOLDEN = os.path.join(ROOT, "tests", "golden")
MPS_FILES = sorted(glob.glob(os.path.join(GOLDEN, "*.mps")) + glob.glob(os.path.join(GOLDEN, "netlib", "*.mps")))
PUBLISHED = json.load(open(os.path.join(GOLDEN, "netlib", "published.json")))["instances"]

needs_ref = pytest.mark.skipif(not O.have_ref(), reason="reference library not built (oracle/_ref)")

QUIRK_MPS = """NAME QUIRKS
* comment
OBJSENSE
    MAX
ROWS
 N obj
 E e1
 L l1
 G g1
 E e2
 E e1
COLUMNS
    x obj 1.5D0 e1 1
    x l1 2 g1 3
    MARKER 'MARKER' 'INTORG'
    y obj 1 e2 1
    MARKER 'MARKER' 'INTEND'
    z e1 1e-12 l1 1
    z nosuchrow 4
    x e1 2
RHS
    rhs obj 7 e1 4
    rhs l1 5
    g1 2
    rhs e2 3
RANGES
    rng e1 2 l1 3
    rng g1 -1
    rng e2 -2
BOUNDS
 UP bnd x -4
 FR bnd y
 FX bnd z 2.5
 LO bnd w 1
 BV bnd y
ENDATA
"""

QUIRK_LP = """\\ corner cases of the LP grammar
MAXIMIZE
 profit: 3 x + 2y - z [ 2 x ^2 + 4 x * y ] / 2
S.T.
 c1: x + y <= 4
 2 x - y >= -2
 r3: -1 <= x - z <= 5.5
 c4: y = 1.5e0
 c5: x + z > 1
 c6: - 2 <= w + x
Bounds
 y <= 10
 -inf <= z <= 3
 x >= -2
 w free
 1 <= v <= INF
General
 v
Binary
 b
End
 ignored: q + r <= 1
"""


def _driver():
    if not os.path.exists(DRIVER):
        pytest.fail(f"{DRIVER} is not built (run __graft_entry__.build())")
    return DRIVER


def _mine(path, kind):
    out = subprocess.run([_driver(), "dump_model", path, kind], capture_output=True, text=True, timeout=120,
                         check=True).stdout
    rec = {}
    for line in out.splitlines():
        key, *vals = line.split()
        rec[key] = vals
    m, n, nnz, mx, ints, qnnz = (int(v) for v in rec["dims"][:6])
    d = dict(m=m, n=n, maximize=bool(mx), integers=ints, q_nnz=qnnz, offset=float.fromhex(rec["dims"][6]),
             col_ptr=np.array([int(v) for v in rec["col_ptr"]], np.int64),
             row_idx=np.array([int(v) for v in rec["row_idx"]], np.int32))
    for k in ("vals", "c", "col_lower", "col_upper", "row_lower", "row_upper"):
        d[k] = np.array([float.fromhex(v) for v in rec[k]])
    return d


def _same(mine, ref):
    for k in ("m", "n", "maximize", "integers", "q_nnz", "offset"):
        assert mine[k] == ref[k], (k, mine[k], ref[k])
    for k in ("col_ptr", "row_idx", "vals", "c", "col_lower", "col_upper", "row_lower", "row_upper"):
        a, b = mine[k], ref[k]
        assert a.shape == b.shape and np.array_equal(a.view(np.int64) if a.dtype == np.float64 else a,
                                                     b.view(np.int64) if b.dtype == np.float64 else b), k


@needs_ref
@pytest.mark.parametrize("path", MPS_FILES, ids=[os.path.basename(p) for p in MPS_FILES])
def test_mps_reader_matches_reference(path):
    _same(_mine(path, "mps"), O.ref_read_mps(path))


@needs_ref
def test_lp_reader_matches_reference():
    path = os.path.join(GOLDEN, "crude_blend.lp")
    _same(_mine(path, "lp"), O.ref_read_lp(path))


@needs_ref
@pytest.mark.parametrize("kind", ["mps", "lp"])
def test_reader_corner_cases_match_reference(tmp_path, kind):
    path = str(tmp_path / f"quirks.{kind}")
    with open(path, "w") as f:
        f.write(QUIRK_MPS if kind == "mps" else QUIRK_LP)
    ref = O.ref_read_mps(path) if kind == "mps" else O.ref_read_lp(path)
    _same(_mine(path, kind), ref)


def test_crude_blend_lp_equals_mps():
    """tests/test_io.cpp:111-150 of the reference: the LP and MPS forms read into the same LP."""
    a = _mine(os.path.join(GOLDEN, "crude_blend.lp"), "lp")
    b = _mine(os.path.join(GOLDEN, "crude_blend.mps"), "mps")
    for k in ("m", "n", "maximize"):
        assert a[k] == b[k]
    for k in ("c", "col_lower", "col_upper", "row_lower", "row_upper", "vals"):
        np.testing.assert_allclose(a[k], b[k], rtol=0, atol=1e-12)


@pytest.mark.parametrize("name", sorted(PUBLISHED))
def test_netlib_dimensions_match_published(name):
    """Published Netlib sizes (rows include the objective row; nonzeros include the
    objective's), as tests/test_io.cpp:238-262 of the reference checks for blend."""
    path = os.path.join(GOLDEN, "afiro.mps") if name == "afiro" else os.path.join(GOLDEN, "netlib", f"{name}.mps")
    d = _mine(path, "mps")
    pub = PUBLISHED[name]
    assert d["m"] + 1 == pub["published_rows"] and d["n"] == pub["published_cols"]
    obj_nnz = int(np.count_nonzero(np.abs(d["c"]) > 1e-12))
    assert len(d["vals"]) + obj_nnz <= pub["published_nonzeros"]   # explicit zeros are dropped on read


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(PUBLISHED))
def test_netlib_published_optimum_fp64(engine, name):
    path = os.path.join(GOLDEN, "afiro.mps") if name == "afiro" else os.path.join(GOLDEN, "netlib", f"{name}.mps")
    lp = lp_cases.read_mps(path)
    opt = PUBLISHED[name]["published_optimal"]
    print(f"[case] netlib {name}", flush=True)
    engine.upload(lp)
    r = engine.solve(vp.default_params(vp.MODE_RESTARTED, tolerance=1e-9, max_iterations=2000000, time_limit=100))
    x, _, y = engine.download()
    assert r.status == 1, r.as_dict()
    k = O.rel_kkt(lp, x, y)
    assert k[0] <= 1e-9 * 1.000001, k
    got = float(np.dot(lp.c, x)) + lp.offset
    assert abs(got - opt) <= 1e-6 * max(1.0, abs(opt)), (got, opt)


@pytest.mark.parametrize("path", MPS_FILES, ids=[os.path.basename(p) for p in MPS_FILES])
def test_python_test_reader_agrees_with_cpp_reader(path):
    """The tests' own MPS reader (tests/lp_cases.py), which feeds the GPU tests, reads the
    same LP as the C++ reader."""
    lp = lp_cases.read_mps(path)
    d = _mine(path, "mps")
    assert (lp.m, lp.n, lp.maximize) == (d["m"], d["n"], d["maximize"])
    assert np.array_equal(lp.col_ptr, d["col_ptr"]) and np.array_equal(lp.row_idx, d["row_idx"])
    for k, v in (("vals", lp.vals), ("c", lp.c), ("col_lower", lp.col_lower), ("col_upper", lp.col_upper),
                 ("row_lower", lp.row_lower), ("row_upper", lp.row_upper)):
        assert np.array_equal(v, d[k]), k


@pytest.mark.skipif(not O.have_ref(), reason="oracle/_ref not built (no reference checkout)")
@pytest.mark.parametrize("seed,m,n", [(0, 12, 30), (2, 1, 9), (3, 9, 1), (5, 60, 150)])
def test_mps_writer_round_trip_through_the_reference_reader(tmp_path, seed, m, n):
    """The test MPS writer (lp_cases.write_mps, used by the end-to-end solve tests) against the
    reference's reader: the same matrix, costs, column bounds, sense and row bounds (a ranged
    row's upper bound, lower + range, to one ulp)."""
    import scipy.sparse as sp
    lp = lp_cases.structured_feasible_lp(seed, m, n)
    path = str(tmp_path / "lp.mps")
    lp_cases.write_mps(lp, path)
    r = O.ref_read_mps(path)
    A = sp.csc_matrix((r["vals"], r["row_idx"], r["col_ptr"]), shape=(r["m"], r["n"])).toarray()
    cap = lambda v: np.where(np.abs(v) >= 1e20, np.sign(v) * 1e20, v)
    assert np.array_equal(A, lp.dense()) and np.array_equal(r["c"], lp.c) and r["maximize"] == lp.maximize
    assert np.array_equal(cap(r["col_lower"]), cap(lp.col_lower)) and np.array_equal(cap(r["col_upper"]), cap(lp.col_upper))
    assert np.array_equal(cap(r["row_lower"]), cap(lp.row_lower))
    np.testing.assert_allclose(cap(r["row_upper"]), cap(lp.row_upper), rtol=1e-15, atol=1e-15)
