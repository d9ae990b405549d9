"""vj_op_diagnostics_static (the diagnostics with their partials in a static device array
instead of an allocated buffer) against vj_op_diagnostics, bit for bit, through the C ABI:
two calls in a row, the second on a smaller input, so that partials left by the first call
in the shared scratch must not leak into the second."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, np.float64)).cuda()


def _inputs(seed, m, n):
    rng = np.random.default_rng(seed)
    lr = rng.uniform(-1.0, 0.0, m)
    ur = lr + rng.uniform(0.0, 2.0, m)
    ur[rng.random(m) < 0.1] = 1e20
    ax = rng.uniform(-2.0, 2.0, m)
    lc = np.where(rng.random(n) < 0.2, -1e20, rng.uniform(-1.0, 0.0, n))
    uc = np.where(rng.random(n) < 0.2, 1e20, rng.uniform(0.5, 2.0, n))
    x = rng.uniform(-1.5, 2.5, n)
    aty = rng.standard_normal(n)
    c = rng.standard_normal(n)
    return [ax, lr, ur, aty, c, x, lc, uc]


def _static(lib, m, n, t, stream):
    out = torch.full((4,), -7.0, dtype=torch.float64, device="cuda")
    rc = lib.vj_op_diagnostics_static(m, n, *[ctypes.c_void_p(a.data_ptr()) for a in t],
                                      ctypes.c_void_p(out.data_ptr()), stream)
    assert rc == 0, lib.vj_last_error()
    torch.cuda.synchronize()
    return out.cpu().numpy()


def test_static_scratch_diagnostics_match_allocating_entry():
    import vajra_pdhg as vp
    lib = vp.lib()
    lib.vj_op_diagnostics_static.argtypes = [ctypes.c_int, ctypes.c_int] + [ctypes.c_void_p] * 10
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    # the first call fills 7,813 blocks of partials, the second (smaller) only 1,172
    for seed, m, n in [(11, 1_000_000, 2_000_000), (12, 300_000, 123_457)]:
        print(f"[case] diagnostics static vs allocating m={m} n={n} seed={seed}", flush=True)
        a = _inputs(seed, m, n)
        t = [_dev(v) for v in a]
        want = vp.op_diagnostics(*t)
        got = _static(lib, m, n, t, stream)
        assert np.array_equal(got, want), (got, want)
        assert want[3] == 0.0 and want[0] > 0.0 and want[1] > 0.0
    # limits: beyond the static scratch is an error with a message, no device work
    big = 32768 * 256 + 1
    rc = lib.vj_op_diagnostics_static(big, 1, *([None] * 8), ctypes.c_void_p(1), stream)
    assert rc == vp.ERR_INVALID_ARGUMENT and b"static scratch" in lib.vj_last_error()
