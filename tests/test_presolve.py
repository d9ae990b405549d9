"""Presolve and postsolve of the C++ API (vajra-opt_amd/host/presolve.cpp) against the
reference's own PresolveEngine (src/presolve/presolve.cpp:10-976) compiled from its sources
(oracle/_ref).

Both sides read the same model file with their own reader, presolve it, and postsolve the same
synthetic reduced solution; tests/cpp/presolve_dump.hpp writes both results as one canonical
text (status and message, the eight reduction counts, the reduced model, the index maps, every
reduction record in stack order, the postsolved primal/dual values, basis and quality, doubles
in hex). The texts must be equal: same reductions in the same order, bit for bit.

Inputs: the Netlib fixtures, afiro, crude_blend (MPS and LP), hand-written models that reach
every reduction type and every infeasible / unbounded exit (tests/golden/presolve/), and
random models built to be rich in singletons, doubletons, fixed, free and empty columns. The
reduction counts of the fixtures are also pinned by expected_counts.json (written from the
reference by tests/golden/presolve/make_expected.py), which needs no reference build."""
import glob
import json
import os
import subprocess

import numpy as np
import pytest

import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "vajra-opt_amd", "_build", "test_indus_api")
GOLDEN = os.path.join(ROOT, "tests", "golden")
PRESOLVE = os.path.join(GOLDEN, "presolve")
FIXTURES = (sorted(glob.glob(os.path.join(GOLDEN, "netlib", "*.mps"))) + sorted(glob.glob(os.path.join(GOLDEN, "*.mps")))
            + [os.path.join(GOLDEN, "crude_blend.lp")] + sorted(glob.glob(os.path.join(PRESOLVE, "*.mps"))))
EXPECTED = json.load(open(os.path.join(PRESOLVE, "expected_counts.json")))["models"]

needs_ref = pytest.mark.skipif(not O.have_ref(), reason="reference library not built (oracle/_ref)")


def ours(path, out):
    if not os.path.exists(DRIVER):
        pytest.fail(f"{DRIVER} is not built (run __graft_entry__.build())")
    kind = "lp" if path.endswith(".lp") else "mps"
    subprocess.run([DRIVER, "presolve_dump", path, kind, out], check=True, timeout=120)
    return open(out).read()


def theirs(path, out):
    O.ref_presolve_dump(path, out)
    return open(out).read()


def _compare(path, tmp_path):
    a = theirs(path, str(tmp_path / "ref.txt"))
    b = ours(path, str(tmp_path / "ours.txt"))
    if a != b:
        la, lb = a.splitlines(), b.splitlines()
        first = next((i for i, (x, y) in enumerate(zip(la, lb)) if x != y), min(len(la), len(lb)))
        pytest.fail(f"{os.path.basename(path)}: dumps differ at line {first}:\n ref: {la[first][:300] if first < len(la) else '<end>'}"
                    f"\n ours: {lb[first][:300] if first < len(lb) else '<end>'}")
    return a


@needs_ref
@pytest.mark.parametrize("path", FIXTURES, ids=lambda p: os.path.relpath(p, GOLDEN))
def test_presolve_postsolve_match_reference(path, tmp_path):
    _compare(path, tmp_path)


def test_fixtures_reach_every_reduction_and_exit():
    """The fixture set exercises all eight reduction types, both non-optimal exits and a
    model that presolve solves completely (checked on the pinned reference counts)."""
    totals = np.zeros(8, int)
    for rec in EXPECTED.values():
        totals += np.array(rec["counts"])
    assert (totals > 0).all(), totals
    assert any(r["infeasible"] for r in EXPECTED.values())
    assert any(r["unbounded"] for r in EXPECTED.values())
    assert any(r.get("reduced") == [0, 0, 0] for r in EXPECTED.values())


@pytest.mark.parametrize("path", FIXTURES, ids=lambda p: os.path.relpath(p, GOLDEN))
def test_presolve_counts_pinned(path, tmp_path):
    """Counts, exit and reduced sizes equal those the reference reported (no reference build
    needed)."""
    exp = EXPECTED[os.path.relpath(path, GOLDEN)]
    text = ours(path, str(tmp_path / "ours.txt"))
    got = {}
    for line in text.splitlines():
        k, _, rest = line.partition(" ")
        v = rest.split()
        if k == "counts":
            got["counts"] = [int(x) for x in v[:8]]
        elif k == "infeasible":
            got["infeasible"], got["unbounded"] = int(v[0]), int(v[2])
        elif k == "message":
            got["message"] = rest
        elif k == "reduced":
            got["reduced"] = [int(v[0]), int(v[1]), int(v[3])]
    assert got == exp


def random_presolve_mps(seed, path, m=60, n=80):
    """A random LP rich in what presolve reduces: columns with 0-3 entries, rows of every
    sense (some ranged), free / fixed / boxed / one-sided columns, coefficients from a small
    set or N(0,1), and right-hand sides taken from a point inside the column bounds, so most
    models stay feasible and the reductions run several passes."""
    rng = np.random.default_rng(seed)
    x0 = rng.uniform(-2, 4, n)
    maximize = seed % 4 == 3
    c = np.where(rng.random(n) < 0.2, 0.0, rng.choice([1.0, -1.0, 2.0, 0.5, -3.0], n))
    kinds = rng.choice(["box", "lo", "up", "free", "fix", "mi"], n, p=[0.35, 0.25, 0.1, 0.12, 0.1, 0.08])
    # a column keeps a finite bound on the side its cost pushes it to (most models bounded)
    push_down = (c > 0) != maximize
    kinds = np.where((c != 0) & push_down & np.isin(kinds, ["up", "free", "mi"]), "lo", kinds)
    kinds = np.where((c != 0) & ~push_down & np.isin(kinds, ["lo", "free"]), "up", kinds)
    lo, up = np.zeros(n), np.zeros(n)
    for j, k in enumerate(kinds):
        a, b = sorted(rng.uniform(-3, 6, 2))
        if k == "box":
            lo[j], up[j] = min(a, x0[j]), max(b, x0[j])
        elif k == "lo":
            lo[j], up[j] = min(a, x0[j]), np.inf
        elif k == "up":
            lo[j], up[j] = min(0.0, x0[j]), max(b, x0[j], 0.0)
        elif k == "free":
            lo[j], up[j] = -np.inf, np.inf
        elif k == "fix":
            lo[j] = up[j] = x0[j]
        else:
            lo[j], up[j] = -np.inf, max(b, x0[j])
    cols = []
    for j in range(n):
        k = rng.choice([0, 1, 2, 3], p=[0.08, 0.3, 0.37, 0.25])
        rows = rng.choice(m, k, replace=False)
        vals = np.where(rng.random(k) < 0.6, rng.choice([1.0, -1.0, 2.0, -2.0, 0.5, 3.0], k), rng.standard_normal(k))
        cols.append((rows, vals))
    act = np.zeros(m)
    for j, (rows, vals) in enumerate(cols):
        act[rows] += vals * x0[j]
    senses = rng.choice(["E", "L", "G", "R"], m, p=[0.45, 0.25, 0.2, 0.1])
    with open(path, "w") as f:
        f.write(f"NAME RP{seed}\n")
        if maximize:
            f.write("OBJSENSE\n    MAX\n")
        f.write("ROWS\n N OBJ\n")
        for i in range(m):
            f.write(f" {'E' if senses[i] == 'R' else senses[i]} R{i}\n")
        f.write("COLUMNS\n")
        for j, (rows, vals) in enumerate(cols):
            f.write(f"    C{j} OBJ {float(c[j])!r}\n")
            for r, v in zip(rows, vals):
                f.write(f"    C{j} R{r} {float(v)!r}\n")
        f.write("RHS\n")
        for i in range(m):
            s = senses[i]
            slack = float(rng.choice([0.0, 0.5, 2.0]))
            rhs = act[i] + (slack if s == "L" else (-slack if s == "G" else 0.0))
            f.write(f"    RHS R{i} {float(rhs)!r}\n")
        if (senses == "R").any():
            f.write("RANGES\n")
            for i in np.nonzero(senses == "R")[0]:
                f.write(f"    RNG R{i} {float(rng.choice([1.0, -1.5, 3.0]))!r}\n")
        f.write("BOUNDS\n")
        for j in range(n):
            if kinds[j] == "free":
                f.write(f" FR BND C{j}\n")
            elif kinds[j] == "fix":
                f.write(f" FX BND C{j} {float(lo[j])!r}\n")
            elif kinds[j] == "mi":
                f.write(f" MI BND C{j}\n UP BND C{j} {float(up[j])!r}\n")
            else:
                f.write(f" LO BND C{j} {float(lo[j])!r}\n")
                if np.isfinite(up[j]):
                    f.write(f" UP BND C{j} {float(up[j])!r}\n")
        f.write("ENDATA\n")


@needs_ref
@pytest.mark.parametrize("seed", range(60))
def test_presolve_random_models_match_reference(seed, tmp_path):
    path = str(tmp_path / f"rp{seed}.mps")
    random_presolve_mps(seed, path, m=40 + seed % 5 * 20, n=60 + seed % 7 * 15)
    _compare(path, tmp_path)


@needs_ref
def test_random_models_reach_the_reductions(tmp_path):
    """The random family does exercise the reductions (summed over the seeds the tests use)."""
    totals = np.zeros(8, int)
    exits = 0
    for seed in range(60):
        path = str(tmp_path / f"rp{seed}.mps")
        random_presolve_mps(seed, path, m=40 + seed % 5 * 20, n=60 + seed % 7 * 15)
        text = theirs(path, str(tmp_path / "ref.txt"))
        for line in text.splitlines():
            if line.startswith("counts"):
                totals += np.array([int(v) for v in line.split()[1:9]])
            if line.startswith("infeasible") and ("infeasible 1" in line or "unbounded 1" in line):
                exits += 1
    print("reduction totals over 60 random models:", totals.tolist(), "non-optimal exits:", exits)
    assert (totals[[0, 1, 2, 3, 5, 6, 7]] > 0).all(), totals
    assert exits <= 30, exits


@needs_ref
@pytest.mark.parametrize("kind", ["structured", "feasible"])
@pytest.mark.parametrize("seed,m,n", [(0, 12, 30), (1, 30, 12), (2, 1, 9), (3, 9, 1), (4, 40, 40), (5, 60, 150),
                                      (6, 150, 60), (7, 3, 3)])
def test_presolve_structured_lps_match_reference(kind, seed, m, n, tmp_path):
    """The structured LPs of the solver tests (empty rows and columns, fixed, free and
    one-sided columns, equality, ranged and one-sided rows, tiny coefficients, both senses;
    feasible or not) written as MPS: presolve and postsolve dumps equal the reference's."""
    import lp_cases
    lp = (lp_cases.structured_lp if kind == "structured" else lp_cases.structured_feasible_lp)(seed, m, n)
    path = str(tmp_path / f"{kind}{seed}.mps")
    lp_cases.write_mps(lp, path)
    _compare(path, tmp_path)
