"""The product's host KKT audits against the reference's own, text for text: for a model read
from an MPS file and a given point, indus::verifier::verify_solution (on the point as given),
indus::Solution::recompute_quality (every SolutionQuality field, the row activities and the
reduced costs it writes) and verify_solution again (on the audited solution) print the same
dump (tests/cpp/audit_dump.hpp, doubles in hex, violation messages and summary included) as
the reference compiled from its sources (src/engine/model.cpp:203-374,
src/engine/verifier.cpp:25-330). These audits decide the status guard and `passed` of every
device solve through indus::solve.

Points: the reference-map CPU PDHG's answer, the same perturbed, with wrong objectives, with
no duals, with NaN entries, and under every status (those that claim a point and those that
do not), on the structured LPs (empty rows and columns, fixed, free and one-sided columns,
ranged and equality rows, tiny coefficients, both senses), afiro and crude_blend. The GPU
variant audits points returned by the device (tests/test_gpu_host_audit.py)."""
import os
import subprocess

import numpy as np
import pytest

import lp_cases
import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "vajra-opt_amd", "_build", "test_indus_api")
STRUCTURED = [(0, 12, 30), (1, 30, 12), (2, 1, 9), (3, 9, 1), (4, 40, 40), (5, 60, 150), (6, 150, 60), (7, 3, 3)]


def write_point(path, status, objective, x, row_dual):
    with open(path, "w") as f:
        f.write(f"status {status}\nobjective {float(objective).hex()}\n")
        f.write("x " + " ".join(float(v).hex() for v in x) + "\n")
        f.write("row_dual " + " ".join(float(v).hex() for v in row_dual) + "\n")


def dumps(mps, point, tol, tmp):
    mine, ref = os.path.join(tmp, "mine.txt"), os.path.join(tmp, "ref.txt")
    subprocess.run([DRIVER, "audit_dump", mps, point, repr(tol), mine], check=True, timeout=120)
    assert O.ref_lib().ref_audit_dump(mps.encode(), point.encode(), O._c_dbl(tol), ref.encode()) == 0
    with open(mine) as a, open(ref) as b:
        return a.read(), b.read()


def points(lp, rng):
    """(status, objective, x, row_dual) cases for lp."""
    r = O.pdhg_solve(lp, O.default_opts(max_iterations=3000, check_frequency=100))
    x, rd = r["x"], r["row_dual"]
    obj = float(np.dot(lp.c, x)) + lp.offset
    noisy_x = x + 1e-3 * rng.standard_normal(lp.n)
    noisy_y = rd + 1e-2 * rng.standard_normal(lp.m)
    nan_x = x.copy()
    nan_x[0] = np.nan
    out = [(1, obj, x, rd), (1, obj + 1.0, x, rd), (1, obj, noisy_x, noisy_y), (6, obj, noisy_x, rd),
           (5, obj * 1.01, x, noisy_y), (1, obj, x, np.zeros(0)), (1, obj, nan_x, rd), (1, obj, x, -rd)]
    out += [(s, obj, x, rd) for s in (0, 2, 3, 7, 9)]
    return out


def lps():
    cases = [(f"structured_{s}_{m}x{n}", lp_cases.structured_lp(s, m, n)) for s, m, n in STRUCTURED[:5]]
    cases += [(f"feasible_{s}_{m}x{n}", lp_cases.structured_feasible_lp(s, m, n)) for s, m, n in STRUCTURED]
    return cases


@pytest.mark.skipif(not O.have_ref(), reason="the reference library (oracle/_ref) is not built here")
@pytest.mark.parametrize("tol", [1e-6, 1e-4])
def test_host_audits_equal_reference(tmp_path, tol):
    if not os.path.exists(DRIVER):
        pytest.fail(f"{DRIVER} is not built")
    rng = np.random.default_rng(42)
    cases = lps()
    for name in ("afiro", "crude_blend"):
        cases.append((name, lp_cases.mps_lp(name)))
    compared = 0
    for name, lp in cases:
        mps = os.path.join(tmp_path, f"{name}.mps")
        if name in ("afiro", "crude_blend"):
            mps = os.path.join(lp_cases.GOLDEN, f"{name}.mps")
        else:
            lp_cases.write_mps(lp, mps)
            lp.offset = 0.0   # MPS carries no objective offset
        for k, (status, obj, x, rd) in enumerate(points(lp, rng)):
            pt = os.path.join(tmp_path, f"{name}_{k}.pt")
            write_point(pt, status, obj, x, rd)
            mine, ref = dumps(mps, pt, tol, str(tmp_path))
            assert mine == ref, (name, k, status)
            compared += 1
    assert compared == len(cases) * 13


PRODUCER_LINES = ("# Solution produced by ", "# Git Commit: ", '  "git_commit": ')


def strip_producer(text):
    return "\n".join(l for l in text.splitlines() if not l.startswith(PRODUCER_LINES))


def write_both(mps, pt, tmp):
    mine = (os.path.join(tmp, "mine.sol"), os.path.join(tmp, "mine.json"))
    ref = (os.path.join(tmp, "ref.sol"), os.path.join(tmp, "ref.json"))
    subprocess.run([DRIVER, "write_files", mps, pt, *mine], check=True, timeout=120)
    assert O.ref_lib().ref_write_files(mps.encode(), pt.encode(), ref[0].encode(), ref[1].encode()) == 0
    return mine, ref


def verify_files(mps, sol, tol, tmp):
    out = os.path.join(tmp, "verify.txt")
    passed = O.ref_lib().ref_verify_files(mps.encode(), sol.encode(), O._c_dbl(tol), out.encode())
    assert passed in (0, 1)
    with open(out) as f:
        return passed, f.read()


@pytest.mark.skipif(not O.have_ref(), reason="the reference library (oracle/_ref) is not built here")
def test_writers_equal_reference_and_parse_in_its_verifier(tmp_path):
    """The .sol and JSON files this build writes (src/io/writer.cpp:85-264) equal the
    reference writer's on the same Solution, line for line apart from the producer and commit
    lines, and the reference's verify_files (its reader and .sol parser,
    src/engine/verifier.cpp:352-668) reads this build's .sol into the same verification as
    its own file's. Points of every status that a .sol names; values to 17 digits."""
    rng = np.random.default_rng(7)
    tmp = str(tmp_path)
    cases = [(n, lp) for n, lp in lps() if n.startswith("feasible")][:5]
    cases.append(("afiro", lp_cases.mps_lp("afiro")))
    checked = 0
    for name, lp in cases:
        mps = os.path.join(lp_cases.GOLDEN, "afiro.mps") if name == "afiro" else os.path.join(tmp, f"{name}.mps")
        if name != "afiro":
            lp_cases.write_mps(lp, mps)
            lp.offset = 0.0
        for k, (status, obj, x, rd) in enumerate(points(lp, rng)):
            if status not in (1, 2, 3, 5) or not np.all(np.isfinite(x)) or rd.size == 0:
                continue   # a .sol names OPTIMAL, FEASIBLE, INFEASIBLE or UNBOUNDED; NaN is no file value
            pt = os.path.join(tmp, "p.pt")
            write_point(pt, status, obj, x, rd)
            mine, ref = write_both(mps, pt, tmp)
            for a, b in zip(mine, ref):
                with open(a) as fa, open(b) as fb:
                    assert strip_producer(fa.read()) == strip_producer(fb.read()), (name, k, a)
            pm, vm = verify_files(mps, mine[0], 1e-6, tmp)
            pr, vr = verify_files(mps, ref[0], 1e-6, tmp)
            assert (pm, vm) == (pr, vr), (name, k)
            for bad in ("Malformed", "Unrecognized", "Could not open", "Duplicate", "Unknown"):
                assert bad not in vm, vm
            checked += 1
    assert checked >= 30
