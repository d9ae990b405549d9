"""TEST INFRASTRUCTURE -- a numpy model of the restarted Halpern PDHG engine in FP64 (the
algorithm of vajra-opt_amd/csrc/vj_engine.hip and vj_kernels.hpp, not the reference's), used
to study the algorithm on the CPU: its setup (Ruiz passes and one Pock-Chambolle pass with
cumulative factors bounded to [1/max_scale, max_scale], bound/objective rescaling whose
norms leave out rows and columns at that bound, 64-step power iteration), the correction
coordinates around an FP64 centre, the reflected Halpern step, and the restart rule and
primal-weight controller of the control kernel, checked every K iterations on the original
LP's rel-KKT, and (refine_every_check) the centre moved at every check with an explicit
Halpern anchor. solve_rounded runs the same algorithm with the low-precision sweeps of the
FP32, BF16 and FP16 engines (values rounded, iterates in FP32), for studies of the
refinement. It reproduces the device's FP64 iteration counts on the LPs it was compared on;
it is not the product and nothing in the product calls it."""
import numpy as np
import scipy.sparse as sp

import oracle as O


def setup(lp, max_scale=1e6, exclude_bounded=True, ruiz_passes=10, pock_chambolle=True):
    A = sp.csc_matrix((lp.vals, lp.row_idx, lp.col_ptr), shape=(lp.m, lp.n)).tocsr()
    m, n = A.shape
    inf = np.inf
    c = (-1.0 if lp.maximize else 1.0) * lp.c
    lc = np.where(lp.col_lower <= -1e20, -inf, lp.col_lower)
    uc = np.where(lp.col_upper >= 1e20, inf, lp.col_upper)
    lr = np.where(lp.row_lower <= -1e20, -inf, lp.row_lower)
    ur = np.where(lp.row_upper >= 1e20, inf, lp.row_upper)
    R, C = np.ones(m), np.ones(n)
    for p in range(ruiz_passes + (1 if pock_chambolle else 0)):
        a = abs(A)
        if p < ruiz_passes:
            rn, cn = a.max(axis=1).toarray().ravel(), a.max(axis=0).toarray().ravel()
        else:
            rn, cn = np.asarray(a.sum(axis=1)).ravel(), np.asarray(a.sum(axis=0)).ravel()
        fr = np.where(rn > 1e-11, 1 / np.sqrt(np.where(rn > 0, rn, 1)), 1.0)
        fc = np.where(cn > 1e-11, 1 / np.sqrt(np.where(cn > 0, cn, 1)), 1.0)
        if max_scale:
            fr = np.clip(R * fr, 1 / max_scale, max_scale) / R
            fc = np.clip(C * fc, 1 / max_scale, max_scale) / C
        A = (sp.diags(fr) @ A @ sp.diags(fc)).tocsr()
        R *= fr
        C *= fc
    cs_, lcs, ucs, lrs, urs = c * C, lc / C, uc / C, lr * R, ur * R
    cm, rm = np.ones(n, bool), np.ones(m, bool)
    if max_scale and exclude_bounded:
        cm = (C < max_scale * (1 - 1e-9)) & (C > (1 + 1e-9) / max_scale)
        rm = (R < max_scale * (1 - 1e-9)) & (R > (1 + 1e-9) / max_scale)
    cn = np.linalg.norm(cs_[cm])
    bmag = np.maximum(np.where(np.isfinite(lrs), np.abs(lrs), 0), np.where(np.isfinite(urs), np.abs(urs), 0))
    bn = np.linalg.norm(bmag[rm])
    cs, bs = 1 / (1 + cn), 1 / (1 + bn)
    omega = (cn * cs) / (bn * bs) if cn * cs > 1e-10 and bn * bs > 1e-10 else 1.0
    v = np.ones(n) / np.sqrt(n)
    lam = 1.0
    for _ in range(64):
        u = A.T @ (A @ v)
        lam = np.linalg.norm(u)
        v = u / lam
    return dict(A=A, AT=A.T.tocsr(), c=cs_ * cs, lc=lcs * bs, uc=ucs * bs, lr=lrs * bs, ur=urs * bs, R=R, C=C, cs=cs,
                bs=bs, omega=omega, norm=np.sqrt(lam))


def solve(lp, S=None, K=64, tol=1e-6, max_iters=200000, rho=1.0, artificial=0.5, smoothing=0.5, integral=0.0,
          sufficient=0.2, necessary=0.8, step=0.95, refine_every_check=False, **setup_kw):
    """Returns status (1 optimal, 6 iteration limit), iterations, restarts and the original-space
    point (x, canonical y) of the last check. refine_every_check: a check that does not restart
    moves the centre to the check point and shifts the iterate and the Halpern anchor (ax, ay,
    zero at a restart) with it; every update adds (1 - lam) times the anchor."""
    S = S or setup(lp, **setup_kw)
    A, AT, c, lc, uc, lr, ur = S["A"], S["AT"], S["c"], S["lc"], S["uc"], S["lr"], S["ur"]
    m, n = A.shape
    eta, omega = step / S["norm"], S["omega"]
    xh, yh = np.clip(0.0, lc, uc), np.zeros(m)

    def recentre(xh, yh):
        ax = A @ xh
        return c - AT @ yh, lc - xh, uc - xh, lr - ax, ur - ax

    dh, lcs, ucs, lrs, urs = recentre(xh, yh)
    x, y = np.zeros(n), np.zeros(m)
    ax, ay = np.zeros(n), np.zeros(m)
    k = iters = restarts = 0
    pw_sum = 0.0
    r0 = r_prev = np.inf
    restarted_prev = True
    while True:
        tau, sig = eta / omega, eta * omega
        for it in range(K):
            lam, pull = (k + it + 1) / (k + it + 2), 1.0 / (k + it + 2)
            xt = np.clip(x - tau * (dh - AT @ y), lcs, ucs)
            if it == 0:
                xq0 = np.dot(xt - x, xt - x)
            if it == K - 1:
                xq1, xq2, xT = np.dot(xt - x, xt - x), np.dot(xt - ax, xt - ax), xt
            xbar = 2 * xt - x
            x = lam * ((1 + rho) * xt - rho * x) + pull * ax
            w = A @ xbar
            full = yh + y
            yt = np.where(sig * (w - lrs) <= full, y + sig * (lrs - w),
                          np.where(sig * (w - urs) > full, y + sig * (urs - w), -yh))
            if it == 0:
                yq0 = np.dot(yt - y, yt - y)
            if it == K - 1:
                yq1, yq2, yT = np.dot(yt - y, yt - y), np.dot(yt - ay, yt - ay), yt
            y = lam * ((1 + rho) * yt - rho * y) + pull * ay
        iters += K
        k += K
        xf, yf = xh + xT, yh + yT
        xo, yo = xf * S["C"] / S["bs"], yf * S["R"] / S["cs"]
        kkt = O.rel_kkt(lp, xo, yo)
        if kkt[0] <= tol or iters >= max_iters:
            return dict(status=1 if kkt[0] <= tol else 6, iters=iters, restarts=restarts, kkt=kkt, x=xo, y=yo)
        r = np.sqrt(omega * xq1 + yq1 / omega)
        rf = np.sqrt(omega * xq0 + yq0 / omega)
        if restarted_prev:
            r0 = r_prev = rf
        restart = r <= sufficient * r0 or (r <= necessary * r0 and r > r_prev) or k >= artificial * iters
        r_prev, restarted_prev = r, False
        if restart:
            restarted_prev, restarts, k = True, restarts + 1, 0
            dx, dy = np.sqrt(xq2), np.sqrt(yq2)
            if dx > 1e-10 and dy > 1e-10:
                if integral == 0.0:
                    omega = np.exp(smoothing * np.log(dy / dx) + (1 - smoothing) * np.log(omega))
                else:   # proportional-integral control of e = log(omega |dx| / |dy|)
                    e = np.log(omega) - np.log(dy / dx)
                    pw_sum += e
                    omega = np.exp(np.log(omega) - (smoothing * e + integral * pw_sum))
            xh, yh = xf, yf
            dh, lcs, ucs, lrs, urs = recentre(xh, yh)
            x, y = np.zeros(n), np.zeros(m)
            ax, ay = np.zeros(n), np.zeros(m)
        elif refine_every_check:
            xh, yh = xf, yf
            dh, lcs, ucs, lrs, urs = recentre(xh, yh)
            x, y, ax, ay = x - xT, y - yT, ax - xT, ay - yT


def _bf16(a):
    b = np.asarray(a, np.float32).view(np.uint32).astype(np.uint64)
    return ((b + 0x7FFF + ((b >> 16) & 1)) & 0xFFFF0000).astype(np.uint32).view(np.float32)


ROUNDING = {"fp32": lambda a: np.asarray(a, np.float32), "bf16": _bf16,
            "fp16": lambda a: np.asarray(a, np.float32).astype(np.float16).astype(np.float32)}


def solve_rounded(lp, S=None, values="fp32", refine_every_check=False, K=64, tol=1e-6, max_iters=200000, rho=1.0,
                  artificial=0.5, smoothing=0.5, sufficient=0.2, necessary=0.8, step=0.95):
    """The engine's low-precision sweeps: matrix values rounded to `values` (FP32, BF16 or FP16),
    iterates, shifted data and sums in FP32, the centre, checks and refinement in FP64; the
    centre moves at restarts, or at every check with the Halpern anchor kept as a vector
    (refine_every_check). No drift restart. Returns status, iterations, restarts and rel-KKT."""
    S = S or setup(lp)
    f32 = np.float32
    A, AT, c, lc, uc, lr, ur = S["A"], S["AT"], S["c"], S["lc"], S["uc"], S["lr"], S["ur"]
    rnd = ROUNDING[values]
    Al = sp.csr_matrix((rnd(A.data), A.indices, A.indptr), shape=A.shape)
    ATl = sp.csr_matrix((rnd(AT.data), AT.indices, AT.indptr), shape=AT.shape)
    m, n = A.shape
    eta, omega = step / S["norm"], S["omega"]
    xh, yh = np.clip(0.0, lc, uc), np.zeros(m)

    def recentre(xh, yh):
        ax = A @ xh
        return tuple(v.astype(f32) for v in (c - AT @ yh, lc - xh, uc - xh, lr - ax, ur - ax))

    dh, lcs, ucs, lrs, urs = recentre(xh, yh)
    yhf = yh.astype(f32)
    x, y = np.zeros(n, f32), np.zeros(m, f32)
    ax, ay = np.zeros(n, f32), np.zeros(m, f32)
    k = iters = restarts = 0
    r0 = r_prev = np.inf
    restarted_prev = True
    while True:
        tau, sig = f32(eta / omega), f32(eta * omega)
        for it in range(K):
            lam, pull = f32((k + it + 1) / (k + it + 2)), f32(1.0 / (k + it + 2))
            xt = np.clip(x - tau * (dh - ATl @ y), lcs, ucs).astype(f32)
            if it == 0:
                xq0 = float(np.sum((xt.astype(float) - x) ** 2))
            if it == K - 1:
                xq1, xT = float(np.sum((xt.astype(float) - x) ** 2)), xt
                xq2 = float(np.sum((xt.astype(float) - ax) ** 2))
            xbar = (f32(2) * xt - x).astype(f32)
            x = (lam * ((1 + rho) * xt - rho * x) + pull * ax).astype(f32)
            w = (Al @ xbar).astype(f32)
            full = yhf + y
            yt = np.where(sig * (w - lrs) <= full, y + sig * (lrs - w),
                          np.where(sig * (w - urs) > full, y + sig * (urs - w), -yhf)).astype(f32)
            if it == 0:
                yq0 = float(np.sum((yt.astype(float) - y) ** 2))
            if it == K - 1:
                yq1, yT = float(np.sum((yt.astype(float) - y) ** 2)), yt
                yq2 = float(np.sum((yt.astype(float) - ay) ** 2))
            y = (lam * ((1 + rho) * yt - rho * y) + pull * ay).astype(f32)
        iters += K
        k += K
        xf, yf = xh + xT.astype(float), yh + yT.astype(float)
        kkt = O.rel_kkt(lp, xf * S["C"] / S["bs"], yf * S["R"] / S["cs"])
        if kkt[0] <= tol or iters >= max_iters:
            return dict(status=1 if kkt[0] <= tol else 6, iters=iters, restarts=restarts, kkt=kkt[0])
        r = np.sqrt(omega * xq1 + yq1 / omega)
        rf = np.sqrt(omega * xq0 + yq0 / omega)
        if restarted_prev:
            r0 = r_prev = rf
        restart = r <= sufficient * r0 or (r <= necessary * r0 and r > r_prev) or k >= artificial * iters
        r_prev, restarted_prev = r, False
        if restart:
            restarted_prev, restarts, k = True, restarts + 1, 0
            dx, dy = np.sqrt(xq2), np.sqrt(yq2)
            if dx > 1e-10 and dy > 1e-10:
                omega = np.exp(smoothing * np.log(dy / dx) + (1 - smoothing) * np.log(omega))
            xh, yh = xf, yf
            dh, lcs, ucs, lrs, urs = recentre(xh, yh)
            yhf = yh.astype(f32)
            x, y = np.zeros(n, f32), np.zeros(m, f32)
            ax, ay = np.zeros(n, f32), np.zeros(m, f32)
        elif refine_every_check:
            xh, yh = xf, yf
            dh, lcs, ucs, lrs, urs = recentre(xh, yh)
            yhf = yh.astype(f32)
            x, y = (x - xT).astype(f32), (y - yT).astype(f32)
            ax, ay = (ax - xT).astype(f32), (ay - yT).astype(f32)

