"""The C++ solver API (include/indus/*.hpp, built into libindus_hip.so) through its test
driver tests/cpp/test_indus_api.cpp.

CPU: the driver's reference test cases (tests/test_pdhg.cpp:54-436 of the reference) in CPU
mode, and the C++ CPU PDHG (indus::pdhg::CpuPdhgSolver) bit for bit against the oracle on
afiro, crude_blend and a random LP written to MPS: same iteration count, status, primal
point, row duals, reduced costs and row activities.
GPU: the driver in device mode (restarted engine, reference-map engine bit-identical to the
C++ CPU PDHG, batch entry point, FP32/BF16 options)."""
import os
import subprocess

import numpy as np
import pytest

import lp_cases
import vajra_pdhg as vp
import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "vajra-opt_amd", "_build", "test_indus_api")
GOLDEN = os.path.join(ROOT, "tests", "golden")


def _driver():
    if not os.path.exists(DRIVER):
        pytest.fail(f"{DRIVER} is not built (run __graft_entry__.build())")
    return DRIVER


def _dump(path, iters, tol, how):
    out = subprocess.run([_driver(), "dump", path, str(iters), repr(tol), how], capture_output=True, text=True,
                         timeout=600, check=True).stdout
    rec = {}
    for line in out.splitlines():
        key, *vals = line.split()
        rec[key] = vals
    arr = {k: np.array([float.fromhex(v) for v in rec[k]]) for k in ("x", "row_dual", "col_dual", "row_value")}
    arr["status"] = int(rec["status"][0])
    arr["iterations"] = int(rec["iterations"][0])
    arr["objective"] = float.fromhex(rec["objective"][0])
    if "diag" in rec:
        d = rec["diag"]
        arr["diag"] = (int(d[0]), int(d[1]), float.fromhex(d[2]), float.fromhex(d[3]), float.fromhex(d[4]))
    return arr


def write_mps(lp, path):
    """Free-format MPS of an LP whose rows are all equalities (repr() floats round-trip)."""
    with open(path, "w") as f:
        f.write("NAME RANDLP\nROWS\n N OBJ\n")
        for i in range(lp.m):
            assert lp.row_lower[i] == lp.row_upper[i]
            f.write(f" E R{i}\n")
        f.write("COLUMNS\n")
        for j in range(lp.n):
            f.write(f"    C{j} OBJ {float(lp.c[j])!r}\n")
            for p in range(lp.col_ptr[j], lp.col_ptr[j + 1]):
                f.write(f"    C{j} R{lp.row_idx[p]} {float(lp.vals[p])!r}\n")
        f.write("RHS\n")
        for i in range(lp.m):
            f.write(f"    RHS R{i} {float(lp.row_lower[i])!r}\n")
        f.write("BOUNDS\n")
        for j in range(lp.n):
            f.write(f" LO BND C{j} {float(lp.col_lower[j])!r}\n UP BND C{j} {float(lp.col_upper[j])!r}\n")
        f.write("ENDATA\n")


def _same(a, b):
    return a.shape == b.shape and np.array_equal(a.view(np.int64), b.view(np.int64))


def test_driver_cpu_mode():
    r = subprocess.run([_driver(), "cpu", GOLDEN], capture_output=True, text=True, timeout=900)
    print(r.stdout[-4000:])
    assert r.returncode == 0 and "ALL PASSED" in r.stdout


@pytest.mark.parametrize("name,iters", [("afiro", 300), ("afiro", 4000), ("crude_blend", 1000)])
def test_cpp_cpu_pdhg_bit_exact_vs_oracle(name, iters):
    path = os.path.join(GOLDEN, f"{name}.mps")
    got = _dump(path, iters, 1e-6, "cpu")
    lp = lp_cases.read_mps(path)
    o = oracle.pdhg_solve(lp, oracle.default_opts(max_iterations=iters, primal_tolerance=1e-6,
                                                  dual_tolerance=1e-6, gap_tolerance=1e-6))
    assert got["iterations"] == o["iterations"]
    assert got["status"] == o["status"]
    assert got["diag"][1] == o["restart_count"]
    assert got["diag"][2] == o["tau"] and got["diag"][3] == o["sigma"]
    assert got["diag"][4] == o["spectral_norm"]
    assert _same(got["x"], o["x"])
    assert _same(got["row_value"], o["row_value"])
    assert got["objective"] == o["objective"]
    if not lp.maximize:
        assert _same(got["row_dual"], o["row_dual"])
        assert _same(got["col_dual"], o["col_dual"])
    else:
        # the reference reports +y for a maximisation (pdhg.cpp:409-411), which its own
        # quality audit reads with the wrong sign; the C++ API reports the canonical -y for
        # both senses and the reduced costs the quality audit derives from it
        assert lp_cases.read_mps(path).maximize
        assert _same(got["row_dual"], -o["row_dual"])
        q = oracle.recompute_quality(lp, o["x"], -o["row_dual"])
        assert _same(got["col_dual"], q["col_dual"])


def test_cpp_cpu_pdhg_bit_exact_random_lp(tmp_path):
    lp = vp.gen_random_lp(0x5EED2026 + 77, 60, 150, 6)
    path = str(tmp_path / "rand.mps")
    write_mps(lp, path)
    got = _dump(path, 2500, 1e-7, "cpu")
    o = oracle.pdhg_solve(lp, oracle.default_opts(max_iterations=2500, primal_tolerance=1e-7,
                                                  dual_tolerance=1e-7, gap_tolerance=1e-7))
    assert got["iterations"] == o["iterations"] and got["status"] == o["status"]
    assert _same(got["x"], o["x"]) and _same(got["row_dual"], o["row_dual"])
    assert _same(got["col_dual"], o["col_dual"])


@pytest.mark.parametrize("kind", ["structured", "feasible"])
@pytest.mark.parametrize("seed,m,n", [(0, 12, 30), (1, 30, 12), (2, 1, 9), (3, 9, 1), (4, 40, 40), (5, 60, 150),
                                      (6, 150, 60), (7, 3, 3)])
def test_cpp_cpu_pdhg_bit_exact_structured(tmp_path, kind, seed, m, n):
    """The C++ CPU PDHG against the oracle, bit for bit, on the structured LPs (empty rows and
    columns, fixed, free and one-sided columns, ranged, equality and one-sided rows, tiny
    coefficients, both senses; feasible or not), read from MPS files on both sides."""
    lp0 = (lp_cases.structured_lp if kind == "structured" else lp_cases.structured_feasible_lp)(seed, m, n)
    path = str(tmp_path / f"{kind}{seed}.mps")
    lp_cases.write_mps(lp0, path)
    lp = lp_cases.read_mps(path)
    got = _dump(path, 1500, 1e-7, "cpu")
    o = oracle.pdhg_solve(lp, oracle.default_opts(max_iterations=1500, primal_tolerance=1e-7,
                                                  dual_tolerance=1e-7, gap_tolerance=1e-7))
    assert got["iterations"] == o["iterations"] and got["status"] == o["status"]
    assert got["diag"][1] == o["restart_count"] and got["diag"][2] == o["tau"]
    assert _same(got["x"], o["x"]) and _same(got["row_value"], o["row_value"])
    assert got["objective"] == o["objective"]
    sign = -1.0 if lp.maximize else 1.0   # the C++ API reports the canonical dual for both senses
    assert _same(got["row_dual"], sign * o["row_dual"])


@pytest.mark.gpu
def test_driver_gpu_mode():
    r = subprocess.run([_driver(), "gpu", GOLDEN], capture_output=True, text=True, timeout=900)
    print(r.stdout[-6000:])
    assert r.returncode == 0 and "ALL PASSED" in r.stdout


@pytest.mark.gpu
def test_cpp_gpu_reference_map_equals_oracle(tmp_path):
    lp = vp.gen_random_lp(0x5EED2026 + 78, 60, 150, 6)
    path = str(tmp_path / "rand.mps")
    write_mps(lp, path)
    got = _dump(path, 1500, 1e-7, "gpu-reference")
    o = oracle.pdhg_solve(lp, oracle.default_opts(max_iterations=1500, primal_tolerance=1e-7,
                                                  dual_tolerance=1e-7, gap_tolerance=1e-7))
    assert got["iterations"] == o["iterations"] and got["status"] == o["status"]
    assert _same(got["x"], o["x"])
    # the C++ layer reports the canonical dual (one convention for both senses); for a
    # minimisation it is the reference's row_dual
    assert _same(got["row_dual"], o["row_dual"])


COMPARE = os.path.join(ROOT, "vajra-opt_amd", "_build", "engine_comparison")
SUITE2_HEADER = ("Platform,Compiler,PrimalTol,DualTol,IterationLimit,TimeLimit_s,Suite,Instance,Rows,Cols,Nonzeros,"
                 "Engine,GpuDevice,PresolveDim,Doubletons,Status,ComputedObj,PublishedObj,RelErr,Iterations")


def _compare(tmp_path, extra=()):
    if not os.path.exists(COMPARE):
        pytest.fail(f"{COMPARE} is not built (run __graft_entry__.build())")
    csv = str(tmp_path / "rows.csv")
    r = subprocess.run([COMPARE, "--models", GOLDEN, "--csv", csv, "--quick", *extra], capture_output=True,
                       text=True, timeout=900)
    print(r.stdout[-3000:])
    assert r.returncode == 0
    import csv as _csv
    with open(csv) as f:
        header = f.readline()
        f.seek(0)
        rows = list(_csv.DictReader(f))
    assert header.startswith(SUITE2_HEADER)
    return rows


def test_engine_comparison_cpu(tmp_path):
    """Suite-2 rows without a device: the speedup is NOT_MEASURED (contract), the device
    engine falls back to the CPU and the restarted engine is not run."""
    rows = _compare(tmp_path)
    assert [r["Engine"] for r in rows] == ["pdhg_cpu", "pdhg_cuda", "pdhg_cuda_restarted"] * 2
    for r in rows:
        assert r["E2ESpeedup"] in ("N/A", "NOT_MEASURED")
    cpu = [r for r in rows if r["Engine"] == "pdhg_cpu"]
    fb = [r for r in rows if r["Engine"] == "pdhg_cuda"]
    # the fallback runs the same CPU map: same iterations and objective
    for a, b in zip(cpu, fb):
        assert a["Iterations"] == b["Iterations"] and a["ComputedObj"] == b["ComputedObj"]


@pytest.mark.gpu
def test_engine_comparison_gpu(tmp_path):
    """With a device: the reference map reproduces the CPU PDHG (same iterations and
    objective, non-zero transfers), which is what the speedup contract asks of the device
    row; the restarted engine passes verification against the published optima."""
    rows = _compare(tmp_path, ["--random", "2000,5000"])
    by = {(r["Instance"], r["Engine"]): r for r in rows}
    for inst in ("afiro", "crude_blend", "random_2000x5000"):
        c, g = by[(inst, "pdhg_cpu")], by[(inst, "pdhg_cuda")]
        assert c["Iterations"] == g["Iterations"] and c["ComputedObj"] == g["ComputedObj"]
        assert int(g["Transfers"]) > 0
    for inst in ("afiro", "crude_blend"):
        assert by[(inst, "pdhg_cuda_restarted")]["VerificationStatus"] == "PASSED"
    assert by[("random_2000x5000", "pdhg_cuda_restarted")]["Status"] == "OPTIMAL"


def test_presolve_handoff_at_scale():
    """Presolve -> postsolve on a 204,000 x 500,000 LP (3M nonzeros) with planted
    reductions: the hash-map passes finish in about a second here (4 s at 1M x 2M), and
    postsolve restores the feasible point exactly (every removed column at its value,
    primal feasibility verified)."""
    r = subprocess.run([_driver(), "presolve_scale", "200000", "500000"], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0
    fields = dict(kv.split("=") for kv in r.stdout.split() if "=" in kv)
    assert float(fields["apply_s"]) < 30.0 and float(fields["postsolve_s"]) < 30.0


def _api_solve(m, n, prec, tol, pre, scaled=False):
    r = subprocess.run([_driver(), "api_solve", str(m), str(n), prec, repr(tol), str(int(pre)), str(int(scaled))],
                       capture_output=True, text=True, timeout=600)
    print(r.stdout)
    assert r.returncode == 0, r.stderr[-2000:]
    line = next(l for l in r.stdout.splitlines() if l.startswith("api_solve"))
    return dict(kv.split("=") for kv in line.split()[1:])


@pytest.mark.gpu
def test_api_solve_with_presolve_at_scale():
    """indus::solve end to end on the device (default presolve on) for a 102,000 x 200,000
    random LP (6 nonzeros per column) with planted reductions: presolve removes them, the
    device solves the reduced LP, postsolve rebuilds the full point, the status guard
    (1e-7 primal audit) and the verifier accept it, and the objective equals the solve
    without presolve to 1e-6."""
    print("[case] api_solve 100000 x 200000 fp64 1e-9 presolve on/off", flush=True)
    a = _api_solve(100000, 200000, "fp64", 1e-9, True)
    b = _api_solve(100000, 200000, "fp64", 1e-9, False)
    assert int(a["status"]) == 1 and int(b["status"]) == 1, (a, b)
    assert int(a["verified"]) == 1 and int(b["verified"]) == 1
    assert int(a["reductions"]) >= 4000 + 2000
    assert int(a["reduced"].split("x")[1]) < 200000
    oa, ob = float(a["objective"]), float(b["objective"])
    assert abs(oa - ob) <= 1e-6 * (1 + abs(ob)), (oa, ob)
    # 1.2M nonzeros: the autotuner times the check and picks a power of two in [32, 256]
    k = int(a["check_interval"])
    assert k in (32, 64, 128, 256), k


_MPS_CASES = [(0, 12, 30), (1, 30, 12), (2, 1, 9), (3, 9, 1), (4, 40, 40), (5, 60, 150), (6, 150, 60), (7, 3, 3)]


@pytest.mark.gpu
@pytest.mark.parametrize("seed,m,n,prec,presolve",
                         [c + (p, 0) for c in _MPS_CASES for p in ("fp64", "fp32")]
                         + [c + ("fp64", 1) for c in _MPS_CASES if c[0] % 2 == 0])
def test_solve_mps_structured_end_to_end(tmp_path, seed, m, n, prec, presolve):
    """indus::solve on the device from a model file (the MPS reader, presolve, the restarted
    engine, postsolve, the quality audit and the status guard of the C++ API) on the feasible
    structured LPs (empty, fixed, free and one-sided rows and columns, ranged and equality
    rows, tiny coefficients, both senses), at rel-KKT 1e-9 so that the reference's 1e-7
    status guard (model.cpp:616-664) accepts the point.

    Without presolve: OPTIMAL, the host's relative KKT error at most 1e-9 and the SciPy-HiGHS
    optimum (MPS carries no objective offset) to 1e-6. With presolve (minimisation LPs: the
    reference's empty-column reduction picks the bound by the raw cost sign also when
    maximising, a quirk kept for parity, which makes maximisation LPs with empty columns
    look unbounded): OPTIMAL and the HiGHS optimum; the postsolved duals follow the
    reference's postsolve, which does not rebuild duals that meet the KKT conditions (its own
    pipeline, run on the 40 x 40 LP, returns a relative KKT error of 0.99), so they are not
    checked here.

    Without presolve the answer is also written by this build's .sol and JSON writers and the
    reference's own verify_files (its MPS reader, .sol parser and verifier,
    src/engine/verifier.cpp:352-668, from oracle/_ref where it is built) passes the .sol."""
    lp = lp_cases.structured_feasible_lp(seed, m, n)
    opt = lp_cases.highs_optimum(lp) - lp.offset
    path = str(tmp_path / f"s{seed}.mps")
    lp_cases.write_mps(lp, path)
    print(f"[case] solve_mps seed={seed} {m}x{n} {prec} presolve={presolve}", flush=True)
    sol, js = str(tmp_path / "answer.sol"), str(tmp_path / "answer.json")
    r = subprocess.run([_driver(), "solve_mps", path, prec, "1e-9", str(presolve), sol, js], capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    kv = dict(t.split("=", 1) for t in r.stdout.splitlines()[0].split()[1:])
    assert int(kv["status"]) == 1, r.stdout
    if not presolve:
        assert float(kv["rel_kkt"]) <= 1e-9 * 1.0001, r.stdout
    got = float(kv["objective"])
    assert abs(got - opt) <= 1e-6 * (1 + abs(opt)), (got, opt, r.stdout)
    if not presolve and oracle.have_ref():
        out = str(tmp_path / "verify.txt")
        passed = oracle.ref_lib().ref_verify_files(path.encode(), sol.encode(), oracle._c_dbl(1e-6), out.encode())
        with open(out) as f:
            report = f.read()
        assert passed == 1, report
        with open(js) as f:
            assert '"verification_status": "PASSED"' in f.read()


def test_options_store_equals_reference():
    """indus::Options (a key -> member binding table here) behaves as the reference's Options
    (src/engine/options.cpp:6-89): the script of tests/cpp/options_probe.hpp (defaults, every
    key and alias, values of another type than the bound member, re-typed keys) gives the same
    members and get_string / get_double / get_int / get_bool answers, text for text, as the
    reference compiled from its sources (tests/golden/options_probe.txt, written by that build
    through tests/golden/make_golden.py; compared live as well where it is built)."""
    mine = subprocess.run([_driver(), "options_probe"], capture_output=True, text=True, timeout=60, check=True).stdout
    with open(os.path.join(GOLDEN, "options_probe.txt")) as f:
        golden = f.read()
    assert mine == golden
    if oracle.have_ref():
        import ctypes
        buf = ctypes.create_string_buffer(1 << 16)
        oracle.ref_lib().ref_options_probe(buf, len(buf))
        assert buf.value.decode() == mine
