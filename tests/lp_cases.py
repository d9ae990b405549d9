"""Known-answer LPs of the reference test suite and seeded workloads, as LP objects.

The small LPs restate tests/test_pdhg.cpp:54-436 of the reference (same data, same
published optima); the MPS fixtures are the reference's own model files copied to
tests/golden/ (data, Netlib/MRPL models) with optima from
SOVEREIGN_SOLVER_BLUEPRINT/test_models/reference.json and apps/benchmark_runner.cpp:707-709.
"""
import os

import numpy as np

import vajra_pdhg as vp

INF = 1e20
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

CFG0 = dict(seed=0x5EED2026, m=2000, n=5000, k=6)


def cfg0_lp():
    return vp.gen_random_lp(CFG0["seed"], CFG0["m"], CFG0["n"], CFG0["k"])


def kat_lps():
    """name -> (LP, expected objective or None, expected x or None)."""
    L = vp.LP.from_dense
    cases = {}
    # test_pdhg.cpp:54-99  min -x1 - 2 x2, x1+x2<=4, x1<=3, x2<=3 -> -7 at (1,3)
    cases["two_var"] = (L([[1, 1], [1, 0], [0, 1]], [-1, -2], [0, 0], [INF, INF], [-INF] * 3, [4, 3, 3]), -7.0,
                        [1.0, 3.0])
    # test_pdhg.cpp:102-142 equality LP -> 13
    cases["equality"] = (L([[1, 1, 1], [2, 1, 0]], [1, 2, 3], [0, 0, 0], [INF] * 3, [6, 5], [6, 5]), 13.0, None)
    # test_pdhg.cpp:145-188 >=, <=, ranged rows -> 4
    cases["ranged"] = (L([[1, 1], [1, -1], [2, 1]], [2, 3], [0, 0], [10, 10], [2, -INF, 1], [INF, 4, 6]), 4.0, None)
    # test_pdhg.cpp:191-231 free variable -> -5 at (-10, 15)
    cases["free_var"] = (L([[1, 1]], [2, 1], [-10, -INF], [10, INF], [5], [5]), -5.0, [-10.0, 15.0])
    # test_pdhg.cpp:234-272 min/max symmetry -> -15 / 15
    cases["min_sym"] = (L([[1, 1]], [-2, -3], [0, 0], [INF, INF], [-INF], [5]), -15.0, None)
    cases["max_sym"] = (L([[1, 1]], [2, 3], [0, 0], [INF, INF], [-INF], [5], maximize=True), 15.0, None)
    # test_pdhg.cpp:405-436 GPU-request LP -> 3
    cases["gpu_fallback"] = (L([[1, 1]], [1, 1], [0, 0], [10, 10], [3], [10]), 3.0, None)
    # test_pdhg.cpp:275-314 presolve LP -> 6 with x3 = 5 (solved here without presolve)
    cases["presolve_lp"] = (L([[1, 1, 0], [0, 1, 0]], [-1, -1, 2], [0, 0, 5], [INF, INF, 5], [-INF, -INF], [4, 10]),
                            6.0, None)
    return cases


def infeasible_lp():
    # test_pdhg.cpp:317-348
    return vp.LP.from_dense([[1, 1], [1, 1]], [1, 1], [0, 0], [10, 10], [-INF, 3], [1, INF])


# --- MPS reader (fixed and free format; ROWS, COLUMNS, RHS, RANGES, BOUNDS) for fixtures ---
def read_mps(path):
    rows = {}          # name -> (index, type)
    row_order = []
    obj_name = None
    cols = {}          # name -> index
    entries = {}       # (row_idx, col_idx) -> value
    cvec = {}
    rhs = {}
    ranges = {}
    bounds = {}
    section = None
    offset = 0.0
    maximize = False
    with open(path) as f:
        for raw in f:
            line = raw.rstrip("\n")
            if not line.strip() or line.startswith("*"):
                continue
            if not line[0].isspace():
                tok = line.split()
                section = tok[0].upper()
                if section == "OBJSENSE" and len(tok) > 1:
                    maximize = tok[1].upper().startswith("MAX")
                continue
            tok = line.split()
            if section == "OBJSENSE":
                maximize = tok[0].upper().startswith("MAX")
            elif section == "ROWS":
                t, name = tok[0].upper(), tok[1]
                if t == "N":
                    if obj_name is None:
                        obj_name = name
                    continue
                rows[name] = (len(row_order), t)
                row_order.append(name)
            elif section == "COLUMNS":
                if "MARKER" in line:
                    continue
                cname = tok[0]
                if cname not in cols:
                    cols[cname] = len(cols)
                j = cols[cname]
                for q in range(1, len(tok) - 1, 2):
                    rname, val = tok[q], float(tok[q + 1])
                    if rname == obj_name:
                        cvec[j] = cvec.get(j, 0.0) + val
                    elif rname in rows:
                        key = (rows[rname][0], j)
                        entries[key] = entries.get(key, 0.0) + val
            elif section == "RHS":
                toks = tok[1:] if len(tok) % 2 == 1 else tok
                for q in range(0, len(toks) - 1, 2):
                    rname, val = toks[q], float(toks[q + 1])
                    if rname == obj_name:
                        offset = -val
                    else:
                        rhs[rname] = val
            elif section == "RANGES":
                toks = tok[1:] if len(tok) % 2 == 1 else tok
                for q in range(0, len(toks) - 1, 2):
                    ranges[toks[q]] = float(toks[q + 1])
            elif section == "BOUNDS":
                bt = tok[0].upper()
                cname = tok[2] if len(tok) >= 3 else tok[1]
                val = float(tok[3]) if len(tok) >= 4 else 0.0
                bounds.setdefault(cname, []).append((bt, val))
    m, n = len(row_order), len(cols)
    lr, ur = np.zeros(m), np.zeros(m)
    for name, (i, t) in rows.items():
        b = rhs.get(name, 0.0)
        if t == "E":
            lo, hi = b, b
            if name in ranges:
                r = ranges[name]
                lo, hi = (b, b + abs(r)) if r >= 0 else (b - abs(r), b)
        elif t == "L":
            lo, hi = -INF, b
            if name in ranges:
                lo = b - abs(ranges[name])
        else:
            lo, hi = b, INF
            if name in ranges:
                hi = b + abs(ranges[name])
        lr[i], ur[i] = lo, hi
    lc, uc = np.zeros(n), np.full(n, INF)
    for cname, bl in bounds.items():
        j = cols[cname]
        for bt, val in bl:
            if bt == "UP":
                uc[j] = val
                if val < 0 and lc[j] == 0.0:
                    lc[j] = -INF
            elif bt == "LO":
                lc[j] = val
            elif bt == "FX":
                lc[j] = uc[j] = val
            elif bt == "FR":
                lc[j], uc[j] = -INF, INF
            elif bt == "MI":
                lc[j] = -INF
            elif bt == "PL":
                uc[j] = INF
    A = np.zeros((m, n))
    for (i, j), v in entries.items():
        A[i, j] = v
    c = np.zeros(n)
    for j, v in cvec.items():
        c[j] = v
    return vp.LP.from_dense(A, c, lc, uc, lr, ur, maximize=maximize, offset=offset,
                            name=os.path.splitext(os.path.basename(path))[0])


MPS_OPTIMA = {"afiro": -464.75314286, "crude_blend": 214.145945946}


def mps_lp(name):
    return read_mps(os.path.join(GOLDEN, f"{name}.mps"))


def structured_lp(seed, m, n):
    """A random LP with the structure the reference's solver treats specially: empty rows
    and columns, fixed, free and one-sided columns, one-sided and equality rows, tiny
    coefficients, a nonzero objective offset and a sense chosen by the seed."""
    rng = np.random.default_rng(1000 + seed)
    A = rng.standard_normal((m, n)) * (rng.random((m, n)) < 0.25)
    if m > 2:
        A[1, :] = 0.0                       # empty row
    if n > 3:
        A[:, 2] = 0.0                       # empty column
    # tiny coefficients just above the reference's triplet drop threshold (kZeroDrop = 1e-11,
    # sparse.cpp:57-60), so the reference's model keeps them as the oracle does
    A[rng.random((m, n)) < 0.03] = 5e-11
    c = rng.standard_normal(n)
    lc, uc = -rng.random(n), 1 + rng.random(n)
    lc[::5] = -1e20                         # free below
    uc[1::6] = 1e20                         # free above
    lc[3::7] = uc[3::7] = 0.5               # fixed
    lr, ur = -np.ones(m) - rng.random(m), np.ones(m) + rng.random(m)
    lr[::4] = -1e20                         # <= rows
    ur[2::5] = 1e20                         # >= rows
    lr[3::6] = ur[3::6]                     # equality rows
    lp = vp.LP.from_dense(A, c, lc, uc, lr, ur)
    lp.maximize = bool(seed % 2)
    lp.offset = float(seed) - 1.5
    return lp


def structured_feasible_lp(seed, m, n):
    """The structure of structured_lp (empty row and column, fixed, free and one-sided
    columns, one-sided, ranged and equality rows, tiny coefficients, offset, both senses) in
    a feasible, bounded LP: the row bounds are built around A x0 for a point x0 inside the
    column bounds, and the cost is c = A^T y0 + z with y0 of the sign each row type allows
    and z of the sign each column's finite bounds allow (zero on free columns), so (y0, z)
    is dual feasible."""
    rng = np.random.default_rng(2000 + seed)
    A = rng.standard_normal((m, n)) * (rng.random((m, n)) < 0.25)
    if m > 2:
        A[1, :] = 0.0
    if n > 3:
        A[:, 2] = 0.0
    A[rng.random((m, n)) < 0.03] = 5e-11
    lc, uc = -rng.random(n), 1 + rng.random(n)
    lc[::5] = -1e20
    uc[1::6] = 1e20
    lc[3::7] = uc[3::7] = 0.5
    x0 = np.where(lc > -1e20, lc, np.where(uc < 1e20, uc - 2.0, rng.standard_normal(n))) + rng.random(n) * 0.5
    x0 = np.minimum(x0, np.where(uc < 1e20, uc, x0))
    x0[3::7] = 0.5
    ax = A @ x0
    lr, ur = ax - rng.random(m), ax + rng.random(m)
    lr[::4] = -1e20
    ur[2::5] = 1e20
    lr[3::6] = ur[3::6] = ax[3::6]
    y0 = rng.standard_normal(m)
    y0 = np.where(lr <= -1e20, -np.abs(y0), np.where(ur >= 1e20, np.abs(y0), y0))
    y0[(lr <= -1e20) & (ur >= 1e20)] = 0.0   # a free row carries no dual
    z = rng.random(n)
    z = np.where((lc > -1e20) & (uc < 1e20), rng.standard_normal(n), np.where(lc > -1e20, z, np.where(uc < 1e20, -z, 0.0)))
    c = A.T @ y0 + z
    maximize = bool(seed % 2)
    lp = vp.LP.from_dense(A, -c if maximize else c, lc, uc, lr, ur)
    lp.maximize = maximize
    lp.offset = float(seed) - 1.5
    return lp


def highs_optimum(lp):
    """Objective of lp by SciPy's HiGHS (an independent solver), or None if it does not
    report an optimum."""
    from scipy.optimize import linprog
    A = lp.dense()
    sgn = -1.0 if lp.maximize else 1.0
    fin = lambda v: None if abs(v) >= 1e20 else float(v)
    bounds = [(fin(l), fin(u)) for l, u in zip(lp.col_lower, lp.col_upper)]
    a_ub, b_ub, a_eq, b_eq = [], [], [], []
    for i in range(lp.m):
        lo, hi = lp.row_lower[i], lp.row_upper[i]
        if lo == hi:
            a_eq.append(A[i]); b_eq.append(lo)
            continue
        if hi < 1e20:
            a_ub.append(A[i]); b_ub.append(hi)
        if lo > -1e20:
            a_ub.append(-A[i]); b_ub.append(-lo)
    r = linprog(sgn * lp.c, A_ub=np.array(a_ub) if a_ub else None, b_ub=b_ub or None,
                A_eq=np.array(a_eq) if a_eq else None, b_eq=b_eq or None, bounds=bounds, method="highs")
    return sgn * r.fun + lp.offset if r.status == 0 else None


def rescaled_lp(lp, seed, decades=3.0):
    """The same LP with row i scaled by r_i and column j by s_j, both 10^U[-decades, decades]:
    A' = R A S, x = S x', c' = S c, column bounds / s_j, row bounds * r_i. The optimal
    objective is unchanged; the coefficient range grows by up to 10^(4 decades)."""
    rng = np.random.default_rng(3000 + seed)
    r = 10.0 ** rng.uniform(-decades, decades, lp.m)
    s = 10.0 ** rng.uniform(-decades, decades, lp.n)
    A = (lp.dense() * r[:, None]) * s[None, :]
    big = 1e20
    div = lambda v, f: np.where(np.abs(v) >= big, v, v / f)
    mul = lambda v, f: np.where(np.abs(v) >= big, v, v * f)
    out = vp.LP.from_dense(A, lp.c * s, div(lp.col_lower, s), div(lp.col_upper, s), mul(lp.row_lower, r),
                           mul(lp.row_upper, r))
    out.maximize = lp.maximize
    out.offset = lp.offset
    return out


def write_mps(lp, path, name="LPCASE"):
    """Free-format MPS of lp (no objective offset: MPS carries none). Rows: equality -> E,
    both sides finite -> G with a RANGE, upper only -> L, lower only -> G, free -> L with a
    1e20 right-hand side. Columns: FR / MI+UP / LO / FX / LO+UP; every column lists its
    objective entry so that empty columns stay in the file."""
    fin = lambda v: abs(v) < 1e20
    A = lp.dense()
    with open(path, "w") as f:
        f.write(f"NAME {name}\n")
        if lp.maximize:
            f.write("OBJSENSE\n    MAX\n")
        f.write("ROWS\n N OBJ\n")
        kinds, rhs, rng = [], [], {}
        for i in range(lp.m):
            lo, hi = lp.row_lower[i], lp.row_upper[i]
            if fin(lo) and fin(hi) and lo == hi:
                kinds.append("E"); rhs.append(lo)
            elif fin(lo) and fin(hi):
                kinds.append("G"); rhs.append(lo); rng[i] = hi - lo
            elif fin(hi):
                kinds.append("L"); rhs.append(hi)
            elif fin(lo):
                kinds.append("G"); rhs.append(lo)
            else:
                kinds.append("L"); rhs.append(1e20)
            f.write(f" {kinds[-1]} R{i}\n")
        f.write("COLUMNS\n")
        for j in range(lp.n):
            f.write(f"    C{j} OBJ {float(lp.c[j])!r}\n")
            for i in np.nonzero(A[:, j])[0]:
                f.write(f"    C{j} R{i} {float(A[i, j])!r}\n")
        f.write("RHS\n")
        for i in range(lp.m):
            f.write(f"    RHS R{i} {float(rhs[i])!r}\n")
        if rng:
            f.write("RANGES\n")
            for i, r in rng.items():
                f.write(f"    RNG R{i} {float(r)!r}\n")
        f.write("BOUNDS\n")
        for j in range(lp.n):
            lo, hi = lp.col_lower[j], lp.col_upper[j]
            if not fin(lo) and not fin(hi):
                f.write(f" FR BND C{j}\n")
            elif not fin(lo):
                f.write(f" MI BND C{j}\n UP BND C{j} {float(hi)!r}\n")
            elif not fin(hi):
                f.write(f" LO BND C{j} {float(lo)!r}\n")
            elif lo == hi:
                f.write(f" FX BND C{j} {float(lo)!r}\n")
            else:
                f.write(f" LO BND C{j} {float(lo)!r}\n UP BND C{j} {float(hi)!r}\n")
        f.write("ENDATA\n")
