"""CPU tests of the C ABI library and the host-side logic (no GPU compute is called):
the library loads, exports every entry point the headers declare, validates arguments,
reports the absence of a device cleanly, and the seeded generators are deterministic."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

import vajra_pdhg as vp

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = [os.path.join(ROOT, "include", h) for h in ("vajra_pdhg.h", "vajra_lpgen.h")]


def declared_functions(path):
    text = open(path).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(vj_[a-z0-9_]+)\s*\(", text)))


def test_every_declared_entry_point_is_exported():
    out = subprocess.run(["nm", "-D", "--defined-only", vp.LIB_PATH], capture_output=True, text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    declared = [f for h in HEADERS for f in declared_functions(h)]
    assert len(declared) >= 20
    missing = [f for f in declared if f not in exported]
    assert not missing, missing
    assert sorted(declared) == sorted(vp.EXPORTED)


def test_params_defaults_follow_reference():
    p = vp.default_params(vp.MODE_REFERENCE)
    # PdhgOptions (pdhg.hpp:17-28) with the Options mapping of pdhg.cpp:58-68
    assert p.max_iterations == 1000000 and p.check_frequency == 100 and p.restart_stagnation_limit == 200
    assert p.step_size_safety == 0.95 and p.tolerance == 1e-6 and p.gap_tolerance == 1e-6 and p.enable_scaling == 1
    q = vp.default_params(vp.MODE_RESTARTED)
    assert q.mode == vp.MODE_RESTARTED and q.check_frequency == 0 and q.use_graph == 1 and q.reflection == 1.0
    # restart rule and primal-weight smoothing of the restarted engine
    assert (q.restart_sufficient, q.restart_necessary, q.restart_artificial) == (0.2, 0.8, 0.5)
    assert q.primal_weight_smoothing == 0.5 and q.primal_weight_integral == 0.0
    assert q.refine_every_check == 0   # per precision: every check with 16-bit values, restarts otherwise
    assert q.stage_threads == 0        # staged kernels: the autotuner times 1,024 and 256 threads
    assert q.power_iterations == 64 and q.step_scale == 0.95
    assert q.restart_drift == 0.0   # 0: the engine's default (2 with FP32 values, 4 with 16-bit values refined
                                    # at every check, off otherwise)


def test_no_device_is_reported_not_faked():
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU is visible")
    info = vp.device_info()
    assert info.available == 0 and b"None" in info.name
    with pytest.raises(vp.VjError) as e:
        vp.Engine(0)
    assert e.value.code == vp.ERR_NO_DEVICE


def test_op_entry_points_validate_arguments():
    L = vp.lib()
    # negative sizes and bad steps are rejected before any device work
    assert L.vj_op_spmv(-1, None, None, None, None, None, 8, 0, 0, None) == vp.ERR_INVALID_ARGUMENT
    assert L.vj_op_dual_update(4, ctypes.c_double(0.0), None, None, None, None, None) == vp.ERR_INVALID_ARGUMENT
    assert L.vj_op_primal_update(4, ctypes.c_double(-1.0), None, None, None, None, None, None, None) == \
        vp.ERR_INVALID_ARGUMENT
    assert L.vj_op_spmv(0, None, None, None, None, None, 8, 0, 0, None) == vp.ERR_INVALID_ARGUMENT
    rp = np.zeros(1, np.int64)
    assert L.vj_op_spmv(0, rp.ctypes.data_as(ctypes.c_void_p), None, None, None, None, 8, 0, 0, None) == vp.OK


def test_random_generator_is_deterministic_and_feasible():
    lp, x0, y0 = vp.gen_random_lp(42, 300, 700, 6, 0.1, 3.0, want_x0=True)
    lp2 = vp.gen_random_lp(42, 300, 700, 6, 0.1, 3.0)
    for a, b in ((lp.vals, lp2.vals), (lp.row_idx, lp2.row_idx), (lp.c, lp2.c), (lp.row_lower, lp2.row_lower)):
        assert np.array_equal(a, b)
    assert lp.nnz == 700 * 6
    for j in range(lp.n):  # distinct ascending rows per column (Floyd sampling)
        r = lp.row_idx[lp.col_ptr[j]:lp.col_ptr[j + 1]]
        assert np.all(np.diff(r) > 0)
    A = lp.dense()
    np.testing.assert_allclose(A @ x0, lp.row_lower, rtol=1e-12, atol=1e-9)   # x0 feasible
    s0 = lp.c - A.T @ y0                                                       # (y0, s0) dual feasible
    assert np.all(s0 >= -1e-9) and np.all(s0 <= 1 + 1e-9)
    assert np.all((x0 >= 0) & (x0 <= 1)) and np.all(lp.col_upper == 10)
    assert vp.gen_random_lp(43, 300, 700).vals[0] != lp.vals[0]


def test_staircase_generator_structure():
    T, rpp, cpp, k, inv = 5, 30, 80, 8, 10
    lp = vp.gen_staircase_lp(7, T, rpp, cpp, k, inv)
    assert (lp.m, lp.n, lp.nnz) == (T * rpp, T * cpp, T * cpp * k + inv * (T - 1))
    A = lp.dense()
    for t in range(T - 1):
        for q in range(inv):
            j = t * cpp + cpp - inv + q
            assert A[(t + 1) * rpp + q, j] == 1.0            # identity coupling into the next period
    blocks = np.zeros_like(A, dtype=bool)
    for t in range(T):
        blocks[t * rpp:(t + 1) * rpp, t * cpp:(t + 1) * cpp] = True
    coupling = (A != 0) & ~blocks
    assert coupling.sum() == inv * (T - 1)
    assert np.all((lp.c >= 1) & (lp.c <= 10)) and np.all((lp.col_upper >= 5) & (lp.col_upper <= 50))


def test_uniform_ints_range():
    v = vp.gen_uniform_ints(3, 5000, 500, 20000)
    assert v.min() >= 500 and v.max() <= 20000 and len(set(v.tolist())) > 4000
    assert np.array_equal(v, vp.gen_uniform_ints(3, 5000, 500, 20000))


def test_lp_csr_matches_dense():
    lp = vp.gen_random_lp(5, 40, 90, 4)
    rp, ci, v = lp.csr()
    A = lp.dense()
    for i in range(lp.m):
        cols = ci[rp[i]:rp[i + 1]]
        assert np.all(np.diff(cols) > 0)
        np.testing.assert_array_equal(v[rp[i]:rp[i + 1]], A[i, cols])


def test_ctypes_mirrors_match_the_c_header(tmp_path):
    """The ctypes structures of vajra_pdhg.py have the size and field offsets that a C
    compiler gives the structures of include/vajra_pdhg.h (a field added on one side only
    would shift every later field)."""
    import shutil
    import subprocess
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    pairs = [("vj_device_info", vp.DeviceInfo), ("vj_params", vp.Params), ("vj_result", vp.Result),
             ("vj_lp", vp.VjLp), ("vj_kernel_stat", vp.KernelStat), ("vj_batch", vp.VjBatch)]
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "vajra_pdhg.h"', 'int main(void) {']
    for cname, py in pairs:
        lines.append(f'    printf("{cname} size %zu\\n", sizeof({cname}));')
        for fname, _ in py._fields_:
            lines.append(f'    printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines += ['    return 0;', '}']
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    subprocess.run([cc, "-std=c11", "-I", os.path.join(root, "include"), str(src), "-o", str(exe)], check=True)
    got = {}
    for line in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.splitlines():
        a, b, v = line.split()
        got[(a, b)] = int(v)
    for cname, py in pairs:
        assert got[(cname, "size")] == ctypes.sizeof(py), cname
        for fname, _ in py._fields_:
            assert got[(cname, fname)] == getattr(py, fname).offset, (cname, fname)
