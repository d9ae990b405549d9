"""Adds advisory objective values from SciPy's HiGHS (a third-party LP solver, used only as a
cross-check) to tests/golden/golden_meta.json. The reference's own dual simplex stops with
NUMERICAL_ERROR on the 2,000 x 5,000 random LP (golden_meta: cfg0_dual_simplex_status = 9),
so this value, together with the KKT certificate of the solution, pins that objective.

    python tests/golden/make_highs_objectives.py
"""
import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
for p in ("vajra-opt_amd", "tests"):
    sys.path.insert(0, os.path.join(ROOT, p))

import numpy as np  # noqa: E402
from scipy.optimize import linprog  # noqa: E402
from scipy.sparse import csc_matrix  # noqa: E402

import lp_cases  # noqa: E402


def highs(lp):
    A = csc_matrix((lp.vals, lp.row_idx, lp.col_ptr), shape=(lp.m, lp.n))
    c = -lp.c if lp.maximize else lp.c
    inf = lambda v: None if abs(v) >= 1e20 else float(v)
    eq = lp.row_lower == lp.row_upper
    A_ub, b_ub = [], []
    rows = A.tocsr()
    for i in range(lp.m):
        if eq[i]:
            continue
        if lp.row_upper[i] < 1e20:
            A_ub.append(rows[i].toarray()[0]); b_ub.append(lp.row_upper[i])
        if lp.row_lower[i] > -1e20:
            A_ub.append(-rows[i].toarray()[0]); b_ub.append(-lp.row_lower[i])
    kw = {}
    if eq.any():
        kw.update(A_eq=rows[np.nonzero(eq)[0]], b_eq=lp.row_lower[eq])
    if A_ub:
        kw.update(A_ub=np.array(A_ub), b_ub=np.array(b_ub))
    r = linprog(c, bounds=[(inf(a), inf(b)) for a, b in zip(lp.col_lower, lp.col_upper)], method="highs",
                options=dict(primal_feasibility_tolerance=1e-10, dual_feasibility_tolerance=1e-10), **kw)
    assert r.status == 0, r.message
    obj = float(r.fun) * (-1 if lp.maximize else 1) + lp.offset
    return obj


def main():
    path = os.path.join(ROOT, "tests", "golden", "golden_meta.json")
    meta = json.load(open(path))
    meta["cfg0_highs_objective"] = highs(lp_cases.cfg0_lp())
    for name in ("afiro", "crude_blend"):
        meta[f"{name}_highs_objective"] = highs(lp_cases.mps_lp(name))
    with open(path, "w") as f:
        json.dump(meta, f, indent=1, sort_keys=True)
    print(json.dumps(meta, indent=1))


if __name__ == "__main__":
    main()
