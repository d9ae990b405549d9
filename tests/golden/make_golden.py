"""Writes the golden fixtures of tests/golden/ from the reference compiled from its own
sources (oracle/_ref/libindus_ref.so, built by `make -C oracle ref` in a container that has
the reference checkout). Run from the repository root:

    python tests/golden/make_golden.py

Fixtures are data only: inputs are regenerated from seeds (their SHA-256 is stored), and
the outputs are the reference's answers on them.
"""
import hashlib
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
for p in ("vajra-opt_amd", "oracle", "tests"):
    sys.path.insert(0, os.path.join(ROOT, p))

import lp_cases  # noqa: E402
import oracle as O  # noqa: E402

OUT = os.path.join(ROOT, "tests", "golden")
TRAJ_CASES = [(1, 10**9), (10, 10**9), (100, 10**9), (1000, 10**9), (700, 100), (3000, 100)]


def lp_sha(lp):
    h = hashlib.sha256()
    for a in (lp.col_ptr, lp.row_idx, lp.vals, lp.c, lp.col_lower, lp.col_upper, lp.row_lower, lp.row_upper):
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def pack(d, prefix):
    out = {}
    for k, v in d.items():
        out[f"{prefix}{k}"] = np.asarray(v)
    return out


def main():
    if not O.have_ref():
        raise SystemExit("oracle/_ref/libindus_ref.so is not built (make -C oracle ref)")
    meta = {}
    cfg0 = lp_cases.cfg0_lp()
    meta["cfg0_sha256"] = lp_sha(cfg0)

    # reference CPU PDHG trajectories on the 2,000 x 5,000 LP (pdhg.cpp:70-440)
    arrs = {}
    for iters, freq in TRAJ_CASES:
        r = O.ref_solve(cfg0, iteration_limit=iters, check_frequency=freq)
        keep = {k: r[k] for k in ("x", "row_dual", "col_dual", "status", "iterations", "tau", "sigma",
                                 "restart_count", "primal_residual", "dual_residual", "duality_gap",
                                 "objective", "spectral_norm")}
        arrs.update(pack(keep, f"cfg0_{iters}_{freq}_"))
    # Ruiz factors and the spectral norm estimate (scaling.cpp:7-153, pdhg.cpp:11-56)
    vals = cfg0.vals.copy()
    c, lr, ur, lc, uc = cfg0.c.copy(), cfg0.row_lower.copy(), cfg0.row_upper.copy(), cfg0.col_lower.copy(), cfg0.col_upper.copy()
    rs, cs = np.zeros(cfg0.m), np.zeros(cfg0.n)
    O.ref_lib().ref_ruiz(cfg0.m, cfg0.n, O._p(cfg0.col_ptr), O._p(cfg0.row_idx), O._p(vals), O._p(c), O._p(lr),
                         O._p(ur), O._p(lc), O._p(uc), 10, O._c_dbl(0.05), O._p(rs), O._p(cs))
    arrs.update(cfg0_ruiz_row_scale=rs, cfg0_ruiz_col_scale=cs, cfg0_ruiz_vals=vals, cfg0_ruiz_c=c)
    arrs["cfg0_spectral_norm_scaled"] = np.asarray(O.ref_lib().ref_spectral_norm(
        cfg0.m, cfg0.n, O._p(cfg0.col_ptr), O._p(cfg0.row_idx), O._p(vals), 25, O._c_dbl(1e-4)))
    np.savez_compressed(os.path.join(OUT, "ref_cfg0.npz"), **arrs)

    # the reference's PDHG known-answer LPs and dual-simplex objectives
    kat = {}
    for name, (lp, obj, xs) in sorted(lp_cases.kat_lps().items()):
        r = O.ref_solve(lp, tolerance=1e-5, iteration_limit=20000)
        s = O.ref_solve(lp, algorithm="dual_simplex", mode=1, presolve=True)
        kat.update(pack({k: r[k] for k in ("x", "row_dual", "status", "iterations", "objective")}, f"{name}_pdhg_"))
        kat[f"{name}_simplex_objective"] = np.asarray(s["objective"])
        kat[f"{name}_simplex_status"] = np.asarray(s["status"])
    inf = lp_cases.infeasible_lp()
    r = O.ref_solve(inf, iteration_limit=500)
    kat.update(pack({k: r[k] for k in ("x", "row_dual", "status", "iterations")}, "infeasible_pdhg_"))
    for name in ("afiro", "crude_blend"):
        lp = lp_cases.mps_lp(name)
        s = O.ref_solve(lp, algorithm="dual_simplex", mode=1, presolve=True)
        kat[f"{name}_simplex_objective"] = np.asarray(s["objective"])
        kat[f"{name}_simplex_status"] = np.asarray(s["status"])
        r = O.ref_solve(lp, iteration_limit=10)
        kat.update(pack({k: r[k] for k in ("x", "row_dual", "status", "iterations")}, f"{name}_pdhg10_"))
        r = O.ref_solve(lp, tolerance=1e-4, iteration_limit=50000)
        kat.update(pack({k: r[k] for k in ("x", "row_dual", "status", "iterations", "objective")}, f"{name}_pdhg_"))
    np.savez_compressed(os.path.join(OUT, "ref_kats.npz"), **kat)

    # the objective oracle for the 2,000 x 5,000 LP: the reference's dual simplex
    s = O.ref_solve(cfg0, algorithm="dual_simplex", mode=1, presolve=True)
    meta["cfg0_dual_simplex_objective"] = s["objective"]
    meta["cfg0_dual_simplex_status"] = s["status"]
    meta["cfg0_dual_simplex_max_primal_violation"] = s["max_primal_violation"]
    meta["cfg0_dual_simplex_max_dual_violation"] = s["max_dual_violation"]
    # the reference's Options store driven by the script of tests/cpp/options_probe.hpp
    buf = O.ctypes.create_string_buffer(1 << 16)
    O.ref_lib().ref_options_probe(buf, len(buf))
    with open(os.path.join(OUT, "options_probe.txt"), "w") as f:
        f.write(buf.value.decode())
    with open(os.path.join(OUT, "golden_meta.json"), "w") as f:
        json.dump(meta, f, indent=1, sort_keys=True)
    print(json.dumps(meta, indent=1))


if __name__ == "__main__":
    main()
