"""Python binding (ctypes) of the MI355X PDHG engine's C ABI (include/vajra_pdhg.h).

Host-side mirror of the reference PDHG interface for tests and benchmarks: an LP is held
in the reference Model layout (A in CSC, include/indus/model.hpp:16-55), the engine is
driven through vj_ctx_* calls, and results come back in the reference's Solution
conventions (unscaled x, row_dual = -y for minimisation, pdhg.cpp:380-418).

The binding loads _build/libvajra_hip.so from this directory and raises ImportError if it
is missing: there is no CPU fallback on this path.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "libvajra_hip.so")
if not os.path.exists(LIB_PATH):
    raise ImportError(f"{LIB_PATH} is not built; run __graft_entry__.build() (make -C vajra-opt_amd)")
_lib = ctypes.CDLL(LIB_PATH)

# ---- constants (include/vajra_pdhg.h) -------------------------------------------------
OK, ERR_INVALID_ARGUMENT, ERR_NO_DEVICE, ERR_ALLOC, ERR_LAUNCH, ERR_NOT_READY, ERR_UNSUPPORTED = range(7)
MODE_REFERENCE, MODE_RESTARTED = 0, 1
PREC_FP64, PREC_FP32, PREC_BF16, PREC_FP16 = 0, 1, 2, 3
STATUS = {0: "NOT_SOLVED", 1: "OPTIMAL", 2: "INFEASIBLE", 3: "UNBOUNDED", 5: "FEASIBLE",
          6: "ITERATION_LIMIT", 7: "TIME_LIMIT", 9: "NUMERICAL_ERROR"}
PRECISIONS = {"fp64": PREC_FP64, "fp32": PREC_FP32, "bf16": PREC_BF16, "fp16": PREC_FP16}

_c_int, _c_i64, _c_dbl, _vp = ctypes.c_int, ctypes.c_int64, ctypes.c_double, ctypes.c_void_p


class DeviceInfo(ctypes.Structure):
    _fields_ = [("available", _c_int), ("device_count", _c_int), ("name", ctypes.c_char * 256),
                ("arch", ctypes.c_char * 64), ("total_bytes", ctypes.c_size_t),
                ("free_bytes", ctypes.c_size_t), ("compute_units", _c_int),
                ("wavefront_size", _c_int), ("major", _c_int), ("minor", _c_int)]


class Params(ctypes.Structure):
    _fields_ = [("mode", _c_int), ("precision", _c_int), ("max_iterations", _c_i64),
                ("time_limit", _c_dbl), ("tolerance", _c_dbl), ("gap_tolerance", _c_dbl),
                ("check_frequency", _c_int), ("restart_stagnation_limit", _c_int),
                ("step_size_safety", _c_dbl), ("enable_scaling", _c_int),
                ("ruiz_iterations", _c_int), ("pock_chambolle", _c_int),
                ("bound_objective_rescaling", _c_int), ("reflection", _c_dbl),
                ("step_scale", _c_dbl), ("use_graph", _c_int), ("lanes_A", _c_int),
                ("lanes_At", _c_int), ("refine_rounds", _c_int), ("verbose", _c_int),
                ("col_splits_A", _c_int), ("col_splits_At", _c_int),
                ("restart_sufficient", _c_dbl), ("restart_necessary", _c_dbl), ("restart_artificial", _c_dbl),
                ("primal_weight_smoothing", _c_dbl), ("power_iterations", _c_int), ("restart_drift", _c_dbl),
                ("primal_weight_integral", _c_dbl), ("refine_every_check", _c_int), ("stage_threads", _c_int)]


class Result(ctypes.Structure):
    _fields_ = [("status", _c_int), ("converged", _c_int), ("iterations", _c_i64),
                ("restart_count", _c_int), ("primal_residual", _c_dbl), ("dual_residual", _c_dbl),
                ("duality_gap", _c_dbl), ("rel_kkt", _c_dbl), ("rel_primal", _c_dbl),
                ("rel_dual", _c_dbl), ("rel_gap", _c_dbl), ("primal_objective", _c_dbl),
                ("dual_objective", _c_dbl), ("spectral_norm", _c_dbl), ("tau", _c_dbl),
                ("sigma", _c_dbl), ("primal_weight", _c_dbl), ("setup_time", _c_dbl),
                ("iteration_time", _c_dbl), ("total_time", _c_dbl), ("upload_time", _c_dbl),
                ("download_time", _c_dbl), ("time_to_target", _c_dbl),
                ("host_device_transfers", _c_i64), ("refinements", _c_int), ("lanes_A", _c_int),
                ("lanes_At", _c_int), ("check_interval", _c_int), ("checks", _c_i64),
                ("col_splits_A", _c_int), ("col_splits_At", _c_int), ("layout_time", _c_dbl)]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_}
        d["status_name"] = STATUS.get(self.status, str(self.status))
        return d


class VjLp(ctypes.Structure):
    _fields_ = [("m", _c_int), ("n", _c_int), ("nnz", _c_i64)] + [
        (f, _vp) for f in ("csr_row_ptr", "csr_col_idx", "csr_vals", "csc_col_ptr", "csc_row_idx",
                           "csc_vals", "c", "col_lower", "col_upper", "row_lower", "row_upper")
    ] + [("maximize", _c_int), ("objective_offset", _c_dbl)]


class VjBatch(ctypes.Structure):
    _fields_ = [("count", _c_int)] + [
        (f, _vp) for f in ("m", "n", "nnz", "csr_row_ptr", "csr_col_idx", "csr_vals", "csc_col_ptr",
                           "csc_row_idx", "csc_vals", "c", "col_lower", "col_upper", "row_lower",
                           "row_upper", "maximize", "objective_offset")]


class KernelStat(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char * 48), ("avg_us", _c_dbl), ("launches", _c_i64),
                ("bytes_per_launch", _c_dbl)]


# exported symbols of include/vajra_pdhg.h and include/vajra_lpgen.h
EXPORTED = [
    "vj_last_error", "vj_device_probe", "vj_params_default", "vj_ctx_create", "vj_ctx_destroy",
    "vj_problem_upload", "vj_params_set", "vj_solve", "vj_download", "vj_reference_iterates",
    "vj_audit_point", "vj_kernel_stats", "vj_benchmark_kernels", "vj_benchmark_blocks", "vj_scaling_factors",
    "vj_batch_upload",
    "vj_batch_solve", "vj_batch_download", "vj_op_spmv", "vj_op_dual_update",
    "vj_op_primal_update", "vj_op_diagnostics", "vj_op_diagnostics_static", "vj_gen_random_lp", "vj_gen_staircase_lp",
    "vj_gen_uniform_ints",
]

_lib.vj_last_error.restype = ctypes.c_char_p
_lib.vj_params_default.restype = None
_lib.vj_params_default.argtypes = [ctypes.POINTER(Params), _c_int]
_lib.vj_gen_random_lp.argtypes = [ctypes.c_uint64, _c_int, _c_int, _c_int, _c_dbl, _c_dbl] + [_vp] * 10
_lib.vj_gen_staircase_lp.argtypes = [ctypes.c_uint64, _c_int, _c_int, _c_int, _c_int, _c_int] + [_vp] * 9
_lib.vj_gen_uniform_ints.argtypes = [ctypes.c_uint64, _c_int, _c_int, _c_int, _vp]
_lib.vj_op_spmv.argtypes = [_c_int, _vp, _vp, _vp, _vp, _vp, _c_int, _c_int, _c_int, _vp]
_lib.vj_op_dual_update.argtypes = [_c_int, _c_dbl, _vp, _vp, _vp, _vp, _vp]
_lib.vj_op_primal_update.argtypes = [_c_int, _c_dbl, _vp, _vp, _vp, _vp, _vp, _vp, _vp]
_lib.vj_op_diagnostics.argtypes = [_c_int, _c_int] + [_vp] * 9 + [_vp]
_lib.vj_op_diagnostics_static.argtypes = [_c_int, _c_int] + [_vp] * 9 + [_vp]
_lib.vj_reference_iterates.argtypes = [_vp, _c_i64, _vp, _vp, _vp]
_lib.vj_audit_point.argtypes = [_vp, _vp, _vp, _vp]
_lib.vj_download.argtypes = [_vp, _vp, _vp, _vp]
_lib.vj_batch_download.argtypes = [_vp, _vp, _vp]
_lib.vj_benchmark_kernels.argtypes = [_vp, _c_int, _vp, _c_int]
_lib.vj_kernel_stats.argtypes = [_vp, _vp, _c_int]
_lib.vj_benchmark_blocks.argtypes = [_vp, _c_int, _c_int, _vp, _vp]
_lib.vj_scaling_factors.argtypes = [_vp, _vp, _vp]


def lib():
    return _lib


class VjError(RuntimeError):
    def __init__(self, code, where):
        msg = _lib.vj_last_error().decode(errors="replace")
        super().__init__(f"{where} failed with code {code}: {msg}")
        self.code = code


def _check(rc, where):
    if rc != OK:
        raise VjError(rc, where)


def _p(a):
    return a.ctypes.data_as(_vp) if a is not None else None


# ---- LP container -----------------------------------------------------------------------
@dataclass
class LP:
    """LP in the reference Model layout: min/max c^T x + offset, row_lower <= A x <= row_upper,
    col_lower <= x <= col_upper; A in CSC with ascending row indices per column."""
    m: int
    n: int
    col_ptr: np.ndarray       # int64 [n+1]
    row_idx: np.ndarray       # int32 [nnz]
    vals: np.ndarray          # float64 [nnz]
    c: np.ndarray
    col_lower: np.ndarray
    col_upper: np.ndarray
    row_lower: np.ndarray
    row_upper: np.ndarray
    maximize: bool = False
    offset: float = 0.0
    name: str = ""
    _csr: tuple | None = field(default=None, repr=False)

    @property
    def nnz(self) -> int:
        return int(self.col_ptr[-1])

    def csr(self):
        """(row_ptr int64, col_idx int32, vals) with ascending columns per row, the order
        SparseMatrixCSC::to_csr produces (sparse.cpp:147-184)."""
        if self._csr is None:
            cols = np.repeat(np.arange(self.n, dtype=np.int32), np.diff(self.col_ptr))
            order = np.argsort(self.row_idx, kind="stable")
            row_ptr = np.zeros(self.m + 1, np.int64)
            np.cumsum(np.bincount(self.row_idx, minlength=self.m), out=row_ptr[1:])
            self._csr = (row_ptr, np.ascontiguousarray(cols[order]), np.ascontiguousarray(self.vals[order]))
        return self._csr

    def dense(self):
        A = np.zeros((self.m, self.n))
        for j in range(self.n):
            for k in range(self.col_ptr[j], self.col_ptr[j + 1]):
                A[self.row_idx[k], j] += self.vals[k]
        return A

    @staticmethod
    def from_dense(A, c, col_lower, col_upper, row_lower, row_upper, maximize=False, offset=0.0, name=""):
        A = np.asarray(A, dtype=np.float64)
        m, n = A.shape
        col_ptr = [0]
        rows, vals = [], []
        for j in range(n):
            nz = np.nonzero(A[:, j])[0]
            rows.extend(nz.tolist())
            vals.extend(A[nz, j].tolist())
            col_ptr.append(len(rows))
        f = lambda v, k: np.ascontiguousarray(np.asarray(v, dtype=np.float64).reshape(k))
        return LP(m, n, np.array(col_ptr, np.int64), np.array(rows, np.int32), np.array(vals, np.float64),
                  f(c, n), f(col_lower, n), f(col_upper, n), f(row_lower, m), f(row_upper, m),
                  bool(maximize), float(offset), name)

    def canonical_c(self):
        return -self.c if self.maximize else self.c.copy()


def gen_random_lp(seed, m, n, k=6, scaled_frac=0.0, log10_range=0.0, want_x0=False):
    """Random feasible, bounded LP (include/vajra_lpgen.h)."""
    nnz = n * k
    cp = np.zeros(n + 1, np.int64)
    ri = np.zeros(nnz, np.int32)
    v = np.zeros(nnz)
    c, lc, uc = np.zeros(n), np.zeros(n), np.zeros(n)
    lr, ur = np.zeros(m), np.zeros(m)
    x0, y0 = np.zeros(n), np.zeros(m)
    rc = _lib.vj_gen_random_lp(seed, m, n, k, scaled_frac, log10_range, _p(cp), _p(ri), _p(v), _p(c), _p(lc),
                               _p(uc), _p(lr), _p(ur), _p(x0), _p(y0))
    if rc != 0:
        raise ValueError("bad generator arguments")
    lp = LP(m, n, cp, ri, v, c, lc, uc, lr, ur, name=f"random_{m}x{n}_s{seed}")
    return (lp, x0, y0) if want_x0 else lp


def gen_staircase_lp(seed, periods, rows_pp=300, cols_pp=800, k=8, inv=100):
    m, n = periods * rows_pp, periods * cols_pp
    nnz = n * k + inv * (periods - 1)
    cp = np.zeros(n + 1, np.int64)
    ri = np.zeros(nnz, np.int32)
    v = np.zeros(nnz)
    c, lc, uc = np.zeros(n), np.zeros(n), np.zeros(n)
    lr, ur = np.zeros(m), np.zeros(m)
    rc = _lib.vj_gen_staircase_lp(seed, periods, rows_pp, cols_pp, k, inv, _p(cp), _p(ri), _p(v), _p(c), _p(lc),
                                  _p(uc), _p(lr), _p(ur), None)
    if rc != 0:
        raise ValueError("bad generator arguments")
    return LP(m, n, cp, ri, v, c, lc, uc, lr, ur, name=f"staircase_T{periods}_s{seed}")


def gen_uniform_ints(seed, count, lo, hi):
    out = np.zeros(count, np.int32)
    _check(_lib.vj_gen_uniform_ints(seed, count, lo, hi, _p(out)), "vj_gen_uniform_ints")
    return out


def device_info():
    d = DeviceInfo()
    _check(_lib.vj_device_probe(ctypes.byref(d)), "vj_device_probe")
    return d


def default_params(mode=MODE_RESTARTED, **kw):
    p = Params()
    _lib.vj_params_default(ctypes.byref(p), mode)
    for k, v in kw.items():
        if k == "precision" and isinstance(v, str):
            v = PRECISIONS[v]
        setattr(p, k, v)
    return p


class Engine:
    """One device context (one HIP stream) holding one uploaded LP or batch."""

    def __init__(self, device=0):
        self._ctx = _vp()
        _check(_lib.vj_ctx_create(device, ctypes.byref(self._ctx)), "vj_ctx_create")
        self.lp = None
        self.batch = None

    def close(self):
        if self._ctx:
            _lib.vj_ctx_destroy(self._ctx)
            self._ctx = _vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- single LP --
    def upload(self, lp: LP, host_csr=False):
        """Upload lp (its CSC). host_csr=True also passes the CSR built on the host; by default
        the engine builds it on the device."""
        rp, ci, av = lp.csr() if host_csr else (None, None, None)
        self._keep = (rp, ci, av)
        s = VjLp(lp.m, lp.n, lp.nnz, _p(rp), _p(ci), _p(av), _p(lp.col_ptr), _p(lp.row_idx), _p(lp.vals),
                 _p(lp.c), _p(lp.col_lower), _p(lp.col_upper), _p(lp.row_lower), _p(lp.row_upper),
                 int(lp.maximize), float(lp.offset))
        _check(_lib.vj_problem_upload(self._ctx, ctypes.byref(s)), "vj_problem_upload")
        self.lp = lp
        self.batch = None

    def set_params(self, p: Params):
        _check(_lib.vj_params_set(self._ctx, ctypes.byref(p)), "vj_params_set")

    def solve(self, params: Params | None = None) -> Result:
        if params is not None:
            self.set_params(params)
        r = Result()
        _check(_lib.vj_solve(self._ctx, ctypes.byref(r)), "vj_solve")
        return r

    def download(self):
        """(x, row_dual in reference convention, canonical y)."""
        x, rd, y = np.zeros(self.lp.n), np.zeros(self.lp.m), np.zeros(self.lp.m)
        _check(_lib.vj_download(self._ctx, _p(x), _p(rd), _p(y)), "vj_download")
        return x, rd, y

    def reference_iterates(self, k):
        x, xb, y = np.zeros(self.lp.n), np.zeros(self.lp.n), np.zeros(self.lp.m)
        _check(_lib.vj_reference_iterates(self._ctx, k, _p(x), _p(xb), _p(y)), "vj_reference_iterates")
        return x, xb, y

    def audit(self, x, y_canonical):
        out = np.zeros(8)
        x = np.ascontiguousarray(x, np.float64)
        y = np.ascontiguousarray(y_canonical, np.float64)
        _check(_lib.vj_audit_point(self._ctx, _p(x), _p(y), _p(out)), "vj_audit_point")
        return out

    def benchmark_kernels(self, reps=20):
        arr = (KernelStat * 8)()
        k = _lib.vj_benchmark_kernels(self._ctx, reps, ctypes.cast(arr, _vp), 8)
        if k < 0:
            raise VjError(-k, "vj_benchmark_kernels")
        return [{"name": arr[i].name.decode(), "avg_us": arr[i].avg_us, "launches": arr[i].launches,
                 "bytes_per_launch": arr[i].bytes_per_launch} for i in range(k)]

    def benchmark_blocks(self, warmup, steps):
        """(ms per check block, iterations per block) of the last restarted solve, replayed on
        its resident problem (vj_benchmark_blocks)."""
        ms = ctypes.c_double()
        k = ctypes.c_int()
        _check(_lib.vj_benchmark_blocks(self._ctx, warmup, steps, ctypes.byref(ms), ctypes.byref(k)),
               "vj_benchmark_blocks")
        return ms.value, k.value

    def scaling_factors(self, m, n):
        """(row_scale[m], col_scale[n]): the device equilibration factors of the last restarted
        solve (vj_scaling_factors)."""
        r, c = np.zeros(m), np.zeros(n)
        _check(_lib.vj_scaling_factors(self._ctx, _p(r), _p(c)), "vj_scaling_factors")
        return r, c

    # -- batches --
    def upload_batch(self, lps, host_csr=False):
        """Upload the LPs as one block-diagonal batch (their CSCs; host_csr=True also passes
        CSRs built on the host, else the engine builds them on the device)."""
        cnt = len(lps)
        ms = np.array([lp.m for lp in lps], np.int32)
        ns = np.array([lp.n for lp in lps], np.int32)
        zs = np.array([lp.nnz for lp in lps], np.int64)
        cat = lambda arrs, dt: np.ascontiguousarray(np.concatenate(arrs).astype(dt, copy=False))
        if host_csr:
            csr = [lp.csr() for lp in lps]
            host = dict(rp=cat([c[0] for c in csr], np.int64), ci=cat([c[1] for c in csr], np.int32),
                        av=cat([c[2] for c in csr], np.float64))
        else:
            host = dict(rp=None, ci=None, av=None)
        parts = dict(
            **host, tp=cat([lp.col_ptr for lp in lps], np.int64),
            ti=cat([lp.row_idx for lp in lps], np.int32), tv=cat([lp.vals for lp in lps], np.float64),
            c=cat([lp.c for lp in lps], np.float64),
            mx=np.array([int(lp.maximize) for lp in lps], np.int32),
            off=np.array([float(lp.offset) for lp in lps], np.float64),
            lc=cat([lp.col_lower for lp in lps], np.float64), uc=cat([lp.col_upper for lp in lps], np.float64),
            lr=cat([lp.row_lower for lp in lps], np.float64), ur=cat([lp.row_upper for lp in lps], np.float64))
        self._keep = (ms, ns, zs, parts)
        b = VjBatch(cnt, _p(ms), _p(ns), _p(zs), _p(parts["rp"]), _p(parts["ci"]), _p(parts["av"]),
                    _p(parts["tp"]), _p(parts["ti"]), _p(parts["tv"]), _p(parts["c"]), _p(parts["lc"]),
                    _p(parts["uc"]), _p(parts["lr"]), _p(parts["ur"]), _p(parts["mx"]), _p(parts["off"]))
        _check(_lib.vj_batch_upload(self._ctx, ctypes.byref(b)), "vj_batch_upload")
        self.batch = lps
        self.lp = None

    def solve_batch(self, params: Params | None = None):
        if params is not None:
            self.set_params(params)
        arr = (Result * len(self.batch))()
        _check(_lib.vj_batch_solve(self._ctx, ctypes.cast(arr, _vp)), "vj_batch_solve")
        return list(arr)

    def download_batch(self):
        ntot = sum(lp.n for lp in self.batch)
        mtot = sum(lp.m for lp in self.batch)
        x, y = np.zeros(ntot), np.zeros(mtot)
        _check(_lib.vj_batch_download(self._ctx, _p(x), _p(y)), "vj_batch_download")
        xs, ys, a, b = [], [], 0, 0
        for lp in self.batch:
            xs.append(x[a:a + lp.n])
            ys.append(y[b:b + lp.m])
            a += lp.n
            b += lp.m
        return xs, ys


# ---- kernel-level operators on torch device tensors --------------------------------------
def _hip_runtimes():
    with open("/proc/self/maps") as f:
        return sorted({line.split()[-1] for line in f if "libamdhip64" in line})


def _require_shared_runtime():
    libs = _hip_runtimes()
    if len(libs) > 1:
        raise RuntimeError("two HIP runtimes are loaded (%s): import torch before vajra_pdhg so that both "
                           "share torch's runtime" % ", ".join(libs))


def _dp(t):
    return _vp(t.data_ptr()) if t is not None else None


def _stream(torch_stream=None):
    import torch
    _require_shared_runtime()
    s = torch_stream if torch_stream is not None else torch.cuda.current_stream()
    return _vp(s.cuda_stream)


def op_spmv(row_ptr, col_idx, vals, x, y, lanes=8, exact_order=False, skip_tiny=False, stream=None):
    rows = row_ptr.numel() - 1
    _check(_lib.vj_op_spmv(rows, _dp(row_ptr), _dp(col_idx), _dp(vals), _dp(x), _dp(y), lanes, int(exact_order),
                           int(skip_tiny), _stream(stream)), "vj_op_spmv")


def op_dual_update(sigma, w, row_lower, row_upper, y, stream=None):
    _check(_lib.vj_op_dual_update(y.numel(), sigma, _dp(w), _dp(row_lower), _dp(row_upper), _dp(y),
                                  _stream(stream)), "vj_op_dual_update")


def op_primal_update(tau, v, c, col_lower, col_upper, x, x_bar, stream=None):
    _check(_lib.vj_op_primal_update(x.numel(), tau, _dp(v), _dp(c), _dp(col_lower), _dp(col_upper), _dp(x),
                                    _dp(x_bar), _stream(stream)), "vj_op_primal_update")


def op_diagnostics(Ax, row_lower, row_upper, Aty, c, x, col_lower, col_upper, stream=None):
    out = np.zeros(4)
    _check(_lib.vj_op_diagnostics(Ax.numel(), x.numel(), _dp(Ax), _dp(row_lower), _dp(row_upper), _dp(Aty), _dp(c),
                                  _dp(x), _dp(col_lower), _dp(col_upper), _p(out), _stream(stream)),
           "vj_op_diagnostics")
    return out


def op_diagnostics_static(Ax, row_lower, row_upper, Aty, c, x, col_lower, col_upper, out4, stream=None):
    """vj_op_diagnostics_static: the same four results written to the device tensor out4,
    asynchronously, with no allocation (partials in a static device array)."""
    _check(_lib.vj_op_diagnostics_static(Ax.numel(), x.numel(), _dp(Ax), _dp(row_lower), _dp(row_upper), _dp(Aty),
                                         _dp(c), _dp(x), _dp(col_lower), _dp(col_upper), _dp(out4), _stream(stream)),
           "vj_op_diagnostics_static")
