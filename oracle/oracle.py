"""TEST INFRASTRUCTURE -- ctypes binding of the CPU oracle (oracle/_build/liboracle.so) and,
when it has been built, of the reference library compiled from its own sources
(oracle/_ref/libindus_ref.so). Only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg import this module, and only to compare against.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "_build", "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libindus_ref.so")

_vp, _c_int, _c_i64, _c_dbl = ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, ctypes.c_double


def _p(a):
    return a.ctypes.data_as(_vp) if a is not None else None


class OrLp(ctypes.Structure):
    _fields_ = [("m", _c_int), ("n", _c_int)] + [
        (f, _vp) for f in ("col_ptr", "row_idx", "vals", "c", "col_lower", "col_upper", "row_lower", "row_upper")
    ] + [("maximize", _c_int), ("offset", _c_dbl)]


class OrOpts(ctypes.Structure):
    _fields_ = [("max_iterations", _c_i64), ("time_limit", _c_dbl), ("primal_tolerance", _c_dbl),
                ("dual_tolerance", _c_dbl), ("gap_tolerance", _c_dbl), ("check_frequency", _c_int),
                ("restart_stagnation_limit", _c_int), ("step_size_safety", _c_dbl), ("enable_scaling", _c_int),
                ("kkt_target", _c_dbl), ("stop_at_iter", _c_i64)]


class OrResult(ctypes.Structure):
    _fields_ = [("status", _c_int), ("converged", _c_int), ("iterations", _c_i64), ("restart_count", _c_int),
                ("primal_residual", _c_dbl), ("dual_residual", _c_dbl), ("duality_gap", _c_dbl),
                ("spectral_norm", _c_dbl), ("tau", _c_dbl), ("sigma", _c_dbl), ("setup_time", _c_dbl),
                ("iteration_time", _c_dbl), ("total_time", _c_dbl), ("objective", _c_dbl),
                ("kkt_first_time", _c_dbl), ("kkt_first_iter", _c_i64), ("kkt_last", _c_dbl),
                ("monitor_time", _c_dbl), ("final_tau", _c_dbl), ("final_sigma", _c_dbl)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            raise ImportError(f"{ORACLE_SO} not built (make -C oracle)")
        _lib = ctypes.CDLL(ORACLE_SO)
        _lib.or_spectral_norm.restype = _c_dbl
    return _lib


def _lp_struct(lp):
    keep = [np.ascontiguousarray(a) for a in (lp.col_ptr, lp.row_idx, lp.vals, lp.c, lp.col_lower, lp.col_upper,
                                               lp.row_lower, lp.row_upper)]
    s = OrLp(lp.m, lp.n, *[_p(a) for a in keep], int(lp.maximize), float(lp.offset))
    return s, keep


def default_opts(**kw):
    o = OrOpts()
    lib().or_default_opts(ctypes.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def pdhg_solve(lp, opts=None, want_traj=False):
    """CpuPdhgSolver::solve restated (pdhg.cpp:70-440). Returns dict with x, row_dual,
    col_dual, row_value, result fields (and traj_* if asked)."""
    o = opts if opts is not None else default_opts()
    s, keep = _lp_struct(lp)
    x, rd, cd, rv = np.zeros(lp.n), np.zeros(lp.m), np.zeros(lp.n), np.zeros(lp.m)
    tx = np.zeros(lp.n) if want_traj else None
    txb = np.zeros(lp.n) if want_traj else None
    ty = np.zeros(lp.m) if want_traj else None
    r = OrResult()
    lib().or_pdhg_solve(ctypes.byref(s), ctypes.byref(o), _p(x), _p(rd), _p(cd), _p(rv), ctypes.byref(r), _p(tx),
                        _p(txb), _p(ty))
    out = dict(r.as_dict(), x=x, row_dual=rd, col_dual=cd, row_value=rv)
    if want_traj:
        out.update(traj_x=tx, traj_xbar=txb, traj_y=ty)
    return out


def csc_multiply(lp, x):
    y = np.zeros(lp.m)
    lib().or_csc_multiply(lp.m, lp.n, _p(lp.col_ptr), _p(lp.row_idx), _p(lp.vals), _p(np.ascontiguousarray(x, float)),
                          _p(y))
    return y


def csc_multiply_transpose(lp, y):
    v = np.zeros(lp.n)
    lib().or_csc_multiply_transpose(lp.m, lp.n, _p(lp.col_ptr), _p(lp.row_idx), _p(lp.vals),
                                    _p(np.ascontiguousarray(y, float)), _p(v))
    return v


def csc_to_csr(lp):
    rp, ci, v = np.zeros(lp.m + 1, np.int64), np.zeros(lp.nnz, np.int32), np.zeros(lp.nnz)
    lib().or_csc_to_csr(lp.m, lp.n, _p(lp.col_ptr), _p(lp.row_idx), _p(lp.vals), _p(rp), _p(ci), _p(v))
    return rp, ci, v


def ruiz(lp, max_iters=10, tol=0.05):
    vals, c = lp.vals.copy(), lp.c.copy()
    lr, ur, lc, uc = lp.row_lower.copy(), lp.row_upper.copy(), lp.col_lower.copy(), lp.col_upper.copy()
    rs, cs = np.ones(lp.m), np.ones(lp.n)
    scaled = lib().or_ruiz(lp.m, lp.n, _p(lp.col_ptr), _p(lp.row_idx), _p(vals), _p(c), _p(lr), _p(ur), _p(lc), _p(uc),
                           max_iters, _c_dbl(tol), _p(rs), _p(cs))
    return dict(scaled=scaled, vals=vals, c=c, row_lower=lr, row_upper=ur, col_lower=lc, col_upper=uc,
                row_scale=rs, col_scale=cs)


def spectral_norm(lp, vals=None, max_iter=25, tol=1e-4):
    v = lp.vals if vals is None else vals
    return lib().or_spectral_norm(lp.m, lp.n, _p(lp.col_ptr), _p(lp.row_idx), _p(np.ascontiguousarray(v)), max_iter,
                                  _c_dbl(tol))


def recompute_quality(lp, x, row_dual):
    s, keep = _lp_struct(lp)
    q = np.zeros(5)
    cd = np.zeros(lp.n)
    lib().or_recompute_quality(ctypes.byref(s), _p(np.ascontiguousarray(x, float)),
                               _p(np.ascontiguousarray(row_dual, float)), _p(cd), _p(q))
    return dict(max_primal_violation=q[0], max_dual_violation=q[1], max_complementarity_violation=q[2],
                is_primal_feasible=bool(q[3]), is_dual_feasible=bool(q[4]), col_dual=cd)


def verify_solution(lp, x, row_dual, objective, tol=1e-6):
    s, keep = _lp_struct(lp)
    v = np.zeros(10)
    rd = None if row_dual is None else np.ascontiguousarray(row_dual, float)
    lib().or_verify_solution(ctypes.byref(s), _p(np.ascontiguousarray(x, float)), _p(rd), _c_dbl(objective),
                             _c_dbl(tol), _p(v))
    keys = ["passed", "max_primal_violation", "max_dual_violation", "max_complementarity_violation",
            "objective_discrepancy", "recomputed_objective", "primal_feasible", "dual_feasible",
            "bounds_feasible", "objective_matches"]
    return {k: (bool(v[i]) if k in ("passed", "primal_feasible", "dual_feasible", "bounds_feasible",
                                    "objective_matches") else float(v[i])) for i, k in enumerate(keys)}


def rel_kkt(lp, x, y_canonical):
    s, keep = _lp_struct(lp)
    out = np.zeros(8)
    lib().or_rel_kkt(ctypes.byref(s), _p(np.ascontiguousarray(x, float)), _p(np.ascontiguousarray(y_canonical, float)),
                     _p(out))
    return out


# ---- the reference library compiled from its own sources ---------------------------------
_ref = None


def have_ref():
    return os.path.exists(REF_SO)


def ref_lib():
    global _ref
    if _ref is None:
        _ref = ctypes.CDLL(REF_SO)
        _ref.ref_spectral_norm.restype = _c_dbl
    return _ref


def ref_solve(lp, algorithm="pdhg_cpu", mode=0, presolve=False, scaling=True, tolerance=1e-6, gap_tolerance=1e-6,
              iteration_limit=1000000, check_frequency=100, use_gpu=False):
    x, rd, cd, rv, s = np.zeros(lp.n), np.zeros(lp.m), np.zeros(lp.n), np.zeros(lp.m), np.zeros(24)
    ref_lib().ref_solve(_c_int(lp.m), _c_int(lp.n), _p(lp.col_ptr), _p(lp.row_idx), _p(lp.vals), _p(lp.c),
                        _p(lp.col_lower), _p(lp.col_upper), _p(lp.row_lower), _p(lp.row_upper), _c_int(int(lp.maximize)),
                        _c_dbl(lp.offset), ctypes.c_char_p(algorithm.encode()), _c_int(mode), _c_int(int(presolve)),
                        _c_int(int(scaling)), _c_dbl(tolerance), _c_dbl(gap_tolerance), _c_i64(iteration_limit),
                        _c_i64(check_frequency), _c_int(int(use_gpu)), _p(x), _p(rd), _p(cd), _p(rv), _p(s))
    names = ["status", "objective", "iterations", "tau", "sigma", "spectral_norm", "restart_count", "primal_residual",
             "dual_residual", "duality_gap", "converged", "max_primal_violation", "max_dual_violation",
             "is_primal_feasible", "is_dual_feasible", "max_complementarity_violation", "best_dual_bound",
             "setup_time", "iteration_time", "total_time", "presolve_total_reductions"]
    d = {k: float(s[i]) for i, k in enumerate(names)}
    d.update(x=x, row_dual=rd, col_dual=cd, row_value=rv)
    return d


def _ref_read(fn, path):
    dims = np.zeros(6, np.int32)
    off = ctypes.c_double(0.0)
    fn(path.encode(), _p(dims), ctypes.byref(off), None, None, None, None, None, None, None, None)
    m, n, nnz, mx, ints, qnnz = [int(v) for v in dims]
    cp, ri, v = np.zeros(n + 1, np.int64), np.zeros(nnz, np.int32), np.zeros(nnz)
    c, lc, uc, lr, ur = np.zeros(n), np.zeros(n), np.zeros(n), np.zeros(m), np.zeros(m)
    fn(path.encode(), _p(dims), ctypes.byref(off), _p(cp), _p(ri), _p(v), _p(c), _p(lc), _p(uc), _p(lr), _p(ur))
    return dict(m=m, n=n, col_ptr=cp, row_idx=ri, vals=v, c=c, col_lower=lc, col_upper=uc, row_lower=lr,
                row_upper=ur, maximize=bool(mx), offset=off.value, integers=ints, q_nnz=qnnz)


def ref_read_mps(path):
    """The reference's MPS reader (src/io/mps_reader.cpp), compiled from its sources."""
    return _ref_read(ref_lib().ref_read_mps, path)


def ref_read_lp(path):
    """The reference's LP reader (src/io/lp_reader.cpp), compiled from its sources."""
    return _ref_read(ref_lib().ref_read_lp, path)


def ref_presolve_dump(path, out_path):
    """The reference's PresolveEngine::apply on a model file read by its own reader, then its
    postsolve of the synthetic reduced solution, written as the canonical text dump of
    tests/cpp/presolve_dump.hpp (src/presolve/presolve.cpp:10-976)."""
    lp = 1 if str(path).endswith(".lp") else 0
    rc = ref_lib().ref_presolve_dump(str(path).encode(), _c_int(lp), str(out_path).encode())
    if rc != 0:
        raise OSError(f"ref_presolve_dump failed on {path}")
