// Kernels of the restarted PDHG engine (single LP and block-diagonal batches).
//
// Row-group SpMV: a 256-thread workgroup covers a tile of R = 256/G consecutive rows;
// each row is handled by G consecutive lanes of a 64-wide wavefront (G = 2..32), which
// stride over the row's nonzeros so that the wavefront reads one contiguous span of the
// value and index streams; the group sum is a fixed xor butterfly (deterministic).
// Epilogues (projection, extrapolation, reflected Halpern step, residual partial sums)
// run in the group's lane 0 right after the reduction, so each iteration is two kernels
// with no intermediate vector round trip through HBM.
#pragma once
#include "vj_device.hpp"
#include "vj_internal.hpp"

namespace vj {

constexpr double kInfBound = 1e20;

// ---------------------------------------------------------------------------------
// Generic FP64 SpMV y = A x over CSR (int32 pointers); used by setup (power iteration).
template <int G>
__global__ __launch_bounds__(kBlock) void spmv_rows_f64(int rows, const int* __restrict__ rp,
                                                        const int* __restrict__ ci,
                                                        const double* __restrict__ av,
                                                        const double* __restrict__ x,
                                                        double* __restrict__ y) {
    const int r = blockIdx.x * (kBlock / G) + threadIdx.x / G;
    const int lane = threadIdx.x % G;
    double acc = 0.0;
    if (r < rows)
        for (int p = rp[r] + lane; p < rp[r + 1]; p += G) acc += av[p] * x[ci[p]];
    acc = group_sum<G>(acc);
    if (r < rows && lane == 0) y[r] = acc;
}

// Per-row max |a| (op 0) or sum |a| (op 1).
template <int G>
__global__ __launch_bounds__(kBlock) void row_abs_reduce(int rows, const int* __restrict__ rp,
                                                         const double* __restrict__ av, int op,
                                                         double* __restrict__ out) {
    const int r = blockIdx.x * (kBlock / G) + threadIdx.x / G;
    const int lane = threadIdx.x % G;
    double acc = 0.0;
    if (r < rows)
        for (int p = rp[r] + lane; p < rp[r + 1]; p += G) {
            const double a = fabs(av[p]);
            acc = op == 0 ? (acc < a ? a : acc) : acc + a;
        }
    acc = op == 0 ? group_max<G>(acc) : group_sum<G>(acc);
    if (r < rows && lane == 0) out[r] = acc;
}

// v_p *= s_own[row] * s_other[idx[p]]  (the same product for A and A^T keeps them exact
// transposes of each other).
template <int G>
__global__ __launch_bounds__(kBlock) void scale_rows(int rows, const int* __restrict__ rp,
                                                     const int* __restrict__ idx,
                                                     double* __restrict__ av,
                                                     const double* __restrict__ s_own,
                                                     const double* __restrict__ s_other) {
    const int r = blockIdx.x * (kBlock / G) + threadIdx.x / G;
    const int lane = threadIdx.x % G;
    if (r >= rows) return;
    const double so = s_own[r];
    for (int p = rp[r] + lane; p < rp[r + 1]; p += G) av[p] *= so * s_other[idx[p]];
}

// factor = norm > 1e-11 ? 1/sqrt(norm) : 1, bounded so that the cumulative scale stays in
// [1/kMaxScale, kMaxScale]; scale *= factor.
// The bound matters for a row (column) whose entries are all tiny next to its bounds (cost):
// unbounded equilibration multiplies its bounds (cost) by up to 1/|a|, e.g. 1e10 for 1e-10
// coefficients, and the bound/objective rescaling that follows then divides every other
// row's bounds (every other cost) by that inflated norm, which stalled the iteration on such
// LPs. Matrices whose rows and columns have norms within 1e+-6 are not affected (the 1M x 2M
// LP's factors lie in [0.007, 5500]).
constexpr double kMaxScale = 1e6;
__global__ void make_scale_step(int len, const double* __restrict__ norm,
                                double* __restrict__ step, double* __restrict__ cum) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= len) return;
    const double v = norm[i];
    const double c0 = cum[i];
    double f = v > 1e-11 ? 1.0 / sqrt(v) : 1.0;
    if (c0 * f > kMaxScale) f = kMaxScale / c0;
    else if (c0 * f < 1.0 / kMaxScale) f = (1.0 / kMaxScale) / c0;
    step[i] = f;
    cum[i] = c0 * f;
}

// Elementwise helpers.
__global__ void fill_f64(int64_t len, double* __restrict__ p, double v) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < len) p[i] = v;
}

// Whether a cumulative equilibration factor sits at the bound of make_scale_step.
__device__ __forceinline__ bool scale_at_bound(double s) {
    return s >= kMaxScale * (1.0 - 1e-9) || s <= (1.0 / kMaxScale) * (1.0 + 1e-9);
}

// Per-element contribution for per-instance sums of squares:
// mode 0: v[i]^2; mode 1: (max finite |lo|,|hi|)^2 (norm of the bound vector).
// skip (may be null): elements whose equilibration factor skip[i] sits at its bound add
// nothing (rows / columns of tiny coefficients only, whose bounds / costs equilibration
// inflated; the bound/objective rescaling and the initial primal weight follow the rest).
template <bool MULTI>
__global__ __launch_bounds__(kBlock) void seg_sumsq(int len, int mode, const double* __restrict__ v,
                                                    const double* __restrict__ lo,
                                                    const double* __restrict__ hi,
                                                    const double* __restrict__ skip,
                                                    const int* __restrict__ inst_of,
                                                    double* __restrict__ part) {
    const int tile = blockIdx.x;
    const int i = tile * kBlock + threadIdx.x;
    double c[1] = {0.0};
    if (i < len && !(skip && scale_at_bound(skip[i]))) {
        if (mode == 0) {
            c[0] = v[i] * v[i];
        } else {
            double b = 0.0;
            if (fabs(lo[i]) < kInfBound) b = fabs(lo[i]);
            if (fabs(hi[i]) < kInfBound && fabs(hi[i]) > b) b = fabs(hi[i]);
            c[0] = b * b;
        }
    }
    const int last = (tile + 1) * kBlock < len ? (tile + 1) * kBlock - 1 : len - 1;
    const int f = MULTI ? inst_of[tile * kBlock] : 0;
    const int l = MULTI ? inst_of[last] : 0;
    const int my = (MULTI && i < len) ? inst_of[i] : 0;
    write_tile_partials<1, MULTI>(c, my, f, l, tile, part, 1);
}

// out[inst] = sum of part column q over the instance's tiles (one block per instance).
__global__ __launch_bounds__(kBlock) void seg_finish(const double* __restrict__ part, int NQ, int q,
                                                     const int* __restrict__ off, int R,
                                                     double* __restrict__ out) {
    __shared__ double lds[kBlock];
    const int inst = blockIdx.x;
    const double s = reduce_instance(part, NQ, q, inst, off[inst], off[inst + 1], R, lds);
    if (threadIdx.x == 0) out[inst] = s;
}

// ---------------------------------------------------------------------------------
// Fused iteration kernels.
//
// Coordinates: the engine iterates on a correction (dx, dy) around an FP64 centre
// (xhat, yhat) that is moved to the current fixed-point estimate at every restart, with
// shifted data dhat = c - A^T yhat, [lc - xhat, uc - xhat], [lr - A xhat, ur - A xhat].
// For FP32/BF16 this is FP64 iterative refinement; for FP64 it is exact re-centring.
// The Halpern anchor is the centre, i.e. zero in correction coordinates.
//
// flags bit 0: first iteration of the block (partial q0 = |T(z)-z|^2 part)
// flags bit 1: last iteration of the block (q1 = |T(z)-z|^2 part, q2 = |T(z)|^2 part,
//              and T(z) + centre is written in FP64 for the check)

// Primal epilogue for column j given acc = (A^T y)_j: projection, extrapolation, reflected
// Halpern step, residual contributions c[] at measured iterations.
// Row-owned operands of the primal update, loaded separately so that the tile-streamed
// kernels can issue them before their gathers (the loads then overlap the SpMV).
template <class T>
struct PrimalOps {
    T xk, dh, lo, hi;
};
// The x iterate at the start of iteration `it`. With reflection rho = 1 the Halpern update
// is x+ = lam (2 xt - x) = lam xbar, exactly (2 xt is exact and both round 2 xt - x once), so
// the step writes only xbar and the next iteration recovers x as lam_prev * xbar, bit for
// bit; lam_prev is the previous iteration's lam (the same expression of s.k + it - 1), 0 at
// the start of an epoch, where the restart kernels have zeroed x and xbar. This saves one
// vector write per primal step. Other rho keep x in its own vector.
// Anchored instances (the FP64 centre moved at a check inside the epoch, s.anchored = 1): the
// Halpern anchor is no longer the origin of the correction coordinates but the vector ax (ay
// for rows), and every update adds (1 - lam) times the anchor, with 1 - lam = 1 / (k + 2)
// formed in FP64 (anchor_pull). The anchor is loaded where it is used (not carried from the
// operand loads through the SpMV), so that it holds no register across the products.
template <class T>
__device__ __forceinline__ T anchor_pull(const InstState& s, int it) {
    return static_cast<T>(1.0 / (static_cast<double>(s.k + it) + 2.0));
}
template <class T>
__device__ __forceinline__ T x_start(int j, const T* __restrict__ x, const T* __restrict__ xbar,
                                     const T* __restrict__ ax, const InstState& s, int it, double rho) {
    if (rho != 1.0) return x[j];
    const double kp = static_cast<double>(s.k + it) - 1.0;
    const T lam_prev = static_cast<T>((kp + 1.0) / (kp + 2.0));
    if (s.anchored) return lam_prev * xbar[j] + anchor_pull<T>(s, it - 1) * ax[j];
    return lam_prev * xbar[j];
}

template <class T>
__device__ __forceinline__ PrimalOps<T> primal_ops(int j, const T* __restrict__ x, const T* __restrict__ xbar,
                                                   const T* __restrict__ ax, const T* __restrict__ dhat,
                                                   const T* __restrict__ lcs, const T* __restrict__ ucs,
                                                   const InstState& s, int it, double rho) {
    return PrimalOps<T>{x_start<T>(j, x, xbar, ax, s, it, rho), dhat[j], lcs[j], ucs[j]};
}

// The projected point xt of column j (the values every primal kernel writes derive from it).
template <class T>
__device__ __forceinline__ T primal_point(T acc, const PrimalOps<T>& o, T tau) {
    T xt = o.xk - tau * (o.dh - acc);
    xt = xt < o.lo ? o.lo : xt;
    xt = xt > o.hi ? o.hi : xt;
    return xt;
}
// Residual contributions of one column at the measured iterations, accumulated into c (one
// column per thread starts from zero, so the sums equal the single terms), and the full FP64
// point at the block's last iteration. c[2] is the squared distance from the anchor (the
// distance the epoch moved, for the primal weight).
template <class T>
__device__ __forceinline__ void primal_residuals(T xt, T xk, T an, int j, int flags, const double* __restrict__ xhat,
                                                 double* __restrict__ xfull, double (&c)[3]) {
    const double d = static_cast<double>(xt) - static_cast<double>(xk);
    if (flags & 1) c[0] += d * d;
    if (flags & 2) {
        const double e = static_cast<double>(xt) - static_cast<double>(an);
        c[1] += d * d;
        c[2] += e * e;
        xfull[j] = xhat[j] + static_cast<double>(xt);
    }
}
template <class T>
__device__ __forceinline__ T anchor_weight(const InstState& s, int it) {
    const double kk = static_cast<double>(s.k + it);
    return static_cast<T>((kk + 1.0) / (kk + 2.0));
}

template <class T>
__device__ __forceinline__ void primal_update(T acc, const PrimalOps<T>& o, int j, const InstState& s, int it,
                                              int flags, double rho, T* __restrict__ x, T* __restrict__ xbar,
                                              const T* __restrict__ ax, const double* __restrict__ xhat,
                                              double* __restrict__ xfull, double (&c)[3]) {
    const T xk = o.xk;
    const T xt = primal_point<T>(acc, o, static_cast<T>(s.tau));
    xbar[j] = static_cast<T>(2) * xt - xk;
    T an = static_cast<T>(0);
    if (s.anchored && (rho != 1.0 || (flags & 2))) an = ax[j];
    if (rho != 1.0) {
        T xn = anchor_weight<T>(s, it) * ((static_cast<T>(1) + static_cast<T>(rho)) * xt - static_cast<T>(rho) * xk);
        if (s.anchored) xn += anchor_pull<T>(s, it) * an;
        x[j] = xn;
    }
    primal_residuals<T>(xt, xk, an, j, flags, xhat, xfull, c);
}

template <class T>
__device__ __forceinline__ void primal_epilogue(T acc, int j, const InstState& s, int it, int flags, double rho,
                                                const T* __restrict__ dhat, const T* __restrict__ lcs,
                                                const T* __restrict__ ucs, T* __restrict__ x, T* __restrict__ xbar,
                                                const T* __restrict__ ax, const double* __restrict__ xhat,
                                                double* __restrict__ xfull, double (&c)[3]) {
    primal_update<T>(acc, primal_ops<T>(j, x, xbar, ax, dhat, lcs, ucs, s, it, rho), j, s, it, flags, rho, x, xbar,
                     ax, xhat, xfull, c);
}

// Dual epilogue for row i given w = (A x_bar)_i: prox of the row-bound support function
// written per active side, y+ = y + sigma*(proj_[l,u](w - y/sigma) - w) in full
// coordinates, then the reflected Halpern step.
// Equality rows: when every row of the problem has l = u, the host passes the same vector
// for both shifted bounds (lrs == urs), and the update needs neither the upper bound nor the
// centre's dual: with l = u the prox is y + sigma (l - w) on every branch below (the first
// branch takes the tie sigma (w - l) = y_full, so no branch depends on y_full), which saves
// two of the four row operands a dual step reads.
template <class T>
struct DualOps {
    T yk, yh, lo, hi;
    bool eq;
};
template <class T>
__device__ __forceinline__ DualOps<T> dual_ops(int i, const T* __restrict__ y, const T* __restrict__ yhat_t,
                                               const T* __restrict__ lrs, const T* __restrict__ urs) {
    if (lrs == urs) {
        const T b = lrs[i];
        return DualOps<T>{y[i], static_cast<T>(0), b, b, true};
    }
    return DualOps<T>{y[i], yhat_t[i], lrs[i], urs[i], false};
}

// The prox point yt of row i (the value the dual kernels write derives from it).
template <class T>
__device__ __forceinline__ T dual_point(T w, const DualOps<T>& o, T sig) {
    const T yk = o.yk;
    if (o.eq) return yk + sig * (o.lo - w);
    const T sfull = o.yh + yk;
    if (sig * (w - o.lo) <= sfull) return yk + sig * (o.lo - w);
    if (sig * (w - o.hi) > sfull) return yk + sig * (o.hi - w);
    return -o.yh;
}
// lam = anchor_weight, pull = anchor_pull of the iteration (pull is used by anchored instances only)
template <class T>
__device__ __forceinline__ T dual_next(T yt, T yk, T an, T lam, T pull, bool anchored, double rho) {
    const T yn = lam * ((static_cast<T>(1) + static_cast<T>(rho)) * yt - static_cast<T>(rho) * yk);
    return anchored ? yn + pull * an : yn;
}
template <class T>
__device__ __forceinline__ void dual_residuals(T yt, T yk, T an, int i, int flags, const double* __restrict__ yhat,
                                               double* __restrict__ yfull, double (&c)[3]) {
    const double d = static_cast<double>(yt) - static_cast<double>(yk);
    if (flags & 1) c[0] += d * d;
    if (flags & 2) {
        const double e = static_cast<double>(yt) - static_cast<double>(an);
        c[1] += d * d;
        c[2] += e * e;
        yfull[i] = yhat[i] + static_cast<double>(yt);
    }
}

template <class T>
__device__ __forceinline__ void dual_update(T w, const DualOps<T>& o, int i, const InstState& s, int it, int flags,
                                            double rho, T* __restrict__ y, const T* __restrict__ ay,
                                            const double* __restrict__ yhat, double* __restrict__ yfull,
                                            double (&c)[3]) {
    const T yt = dual_point<T>(w, o, static_cast<T>(s.sigma));
    const T an = s.anchored ? ay[i] : static_cast<T>(0);
    y[i] = dual_next<T>(yt, o.yk, an, anchor_weight<T>(s, it), anchor_pull<T>(s, it), s.anchored != 0, rho);
    dual_residuals<T>(yt, o.yk, an, i, flags, yhat, yfull, c);
}

template <class T>
__device__ __forceinline__ void dual_epilogue(T w, int i, const InstState& s, int it, int flags, double rho,
                                              const T* __restrict__ yhat_t, const T* __restrict__ lrs,
                                              const T* __restrict__ urs, T* __restrict__ y,
                                              const T* __restrict__ ay, const double* __restrict__ yhat,
                                              double* __restrict__ yfull, double (&c)[3]) {
    dual_update<T>(w, dual_ops<T>(i, y, yhat_t, lrs, urs), i, s, it, flags, rho, y, ay, yhat, yfull, c);
}

template <class V, class T, int G, bool MULTI>
__global__ __launch_bounds__(kBlock) void primal_step(
    int n, const int* __restrict__ tp, const int* __restrict__ ti, const V* __restrict__ tv,
    const T* __restrict__ y, const T* __restrict__ dhat, const T* __restrict__ lcs,
    const T* __restrict__ ucs, T* __restrict__ x, T* __restrict__ xbar, const T* __restrict__ ax,
    const double* __restrict__ xhat, double* __restrict__ xfull,
    const InstState* __restrict__ st, const int* __restrict__ col_inst, double* __restrict__ part,
    int it, int flags, double rho) {
    constexpr int R = kBlock / G;
    const int tile = blockIdx.x;
    const int r0 = tile * R;
    const int j = r0 + static_cast<int>(threadIdx.x) / G;
    const int lane = threadIdx.x % G;
    int f = 0, l = 0;
    if (MULTI) {
        f = col_inst[r0];
        l = col_inst[(r0 + R < n ? r0 + R : n) - 1];
    }
    bool any = false;
    for (int q = f; q <= l; ++q) any |= (st[q].active != 0);
    if (!any) return;
    const int my = (MULTI && j < n) ? col_inst[j] : 0;
    const bool act = (j < n) && (!MULTI || st[my].active != 0);
    double c[3] = {0.0, 0.0, 0.0};
    T acc = 0;
    if (act) {
        const int p1 = tp[j + 1];
        for (int p = tp[j] + lane; p < p1; p += G) acc += static_cast<T>(load_val(tv, p)) * y[load_idx(ti, p)];
    }
    acc = group_sum<G>(acc);
    if (act && lane == 0) primal_epilogue<T>(acc, j, st[my], it, flags, rho, dhat, lcs, ucs, x, xbar, ax, xhat, xfull, c);
    if (flags) {
        const int mask = ((flags & 1) ? 1 : 0) | ((flags & 2) ? 6 : 0);
        write_tile_partials<3, MULTI>(c, my, f, l, tile, part, mask);
    }
}

template <class V, class T, int G, bool MULTI>
__global__ __launch_bounds__(kBlock) void dual_step(
    int m, const int* __restrict__ rp, const int* __restrict__ ci, const V* __restrict__ av,
    const T* __restrict__ xbar, const T* __restrict__ yhat_t, const T* __restrict__ lrs,
    const T* __restrict__ urs, T* __restrict__ y, const T* __restrict__ ay, const double* __restrict__ yhat,
    double* __restrict__ yfull, const InstState* __restrict__ st, const int* __restrict__ row_inst,
    double* __restrict__ part, int it, int flags, double rho) {
    constexpr int R = kBlock / G;
    const int tile = blockIdx.x;
    const int r0 = tile * R;
    const int i = r0 + static_cast<int>(threadIdx.x) / G;
    const int lane = threadIdx.x % G;
    int f = 0, l = 0;
    if (MULTI) {
        f = row_inst[r0];
        l = row_inst[(r0 + R < m ? r0 + R : m) - 1];
    }
    bool any = false;
    for (int q = f; q <= l; ++q) any |= (st[q].active != 0);
    if (!any) return;
    const int my = (MULTI && i < m) ? row_inst[i] : 0;
    const bool act = (i < m) && (!MULTI || st[my].active != 0);
    double c[3] = {0.0, 0.0, 0.0};
    T w = 0;
    if (act) {
        const int p1 = rp[i + 1];
        for (int p = rp[i] + lane; p < p1; p += G) w += static_cast<T>(load_val(av, p)) * xbar[load_idx(ci, p)];
    }
    w = group_sum<G>(w);
    if (act && lane == 0) dual_epilogue<T>(w, i, st[my], it, flags, rho, yhat_t, lrs, urs, y, ay, yhat, yfull, c);
    if (flags) {
        const int mask = ((flags & 1) ? 1 : 0) | ((flags & 2) ? 6 : 0);
        write_tile_partials<3, MULTI>(c, my, f, l, tile, part, mask);
    }
}

// ---------------------------------------------------------------------------------
// Tile-streamed fused steps. The host partitions the rows into tiles of at most 256 rows
// whose nonzeros fit kStreamCap entries (a longer row forms a tile of its own). A workgroup
// streams its tile's contiguous span of values and indices with coalesced loads that do
// not depend on the row structure (many independent loads in flight per lane), writes the
// products to LDS, then one lane per row adds its segment in index order and runs the
// epilogue. A single-row tile longer than kStreamCap is reduced by the whole workgroup.
constexpr int kStreamCap = 4096;

// Entry streams of a CSR matrix as the tile-streamed kernels read them: separate index and
// value arrays, or (FP32 values) interleaved 8-byte (index, value) records, one load per
// entry instead of two.
template <class V>
struct SepEntries {
    const int* __restrict__ idx;
    const V* __restrict__ val;
    template <class T>
    __device__ __forceinline__ void load(int64_t k, int& c, T& a) const {
        c = load_idx(idx, k);
        a = static_cast<T>(load_val(val, k));
    }
};
typedef int vj_i32x2 __attribute__((ext_vector_type(2)));
struct RecEntries {
    const vj_i32x2* __restrict__ rec;
    template <class T>
    __device__ __forceinline__ void load(int64_t k, int& c, T& a) const {
        const vj_i32x2 r = __builtin_nontemporal_load(rec + k);
        c = r.x;
        a = static_cast<T>(__int_as_float(r.y));
    }
};

// Entries of the LDS-staged kernels with 16-bit column offsets from their tile's window start
// (windows up to 65,536 columns): 2 bytes of index per entry instead of 4. The kernels
// subtract the window start from their chunk bases instead (window_relative below).
template <class V>
struct Win16Entries {
    const uint16_t* __restrict__ off;
    const V* __restrict__ val;
    template <class T>
    __device__ __forceinline__ void load(int64_t k, int& c, T& a) const {
        c = static_cast<int>(__builtin_nontemporal_load(off + k));
        a = static_cast<T>(load_val(val, k));
    }
};
// 16-bit values (BF16 or FP16): the 16-bit window offset (low half) and the value's bits (high
// half) in one 4-byte record, one load per entry. For BF16 the high half is the upper half of
// the value's FP32 bit pattern.
template <class V>
struct Win16Records {
    const uint32_t* __restrict__ rec;
    template <class T>
    __device__ __forceinline__ void load(int64_t k, int& c, T& a) const {
        const uint32_t r = __builtin_nontemporal_load(rec + k);
        c = static_cast<int>(r & 0xffffu);
        if constexpr (std::is_same<V, bf16_t>::value)
            a = static_cast<T>(__uint_as_float(r & 0xffff0000u));
        else
            a = static_cast<T>(f16_bits_to_float(static_cast<uint16_t>(r >> 16)));
    }
};
template <class Ent>
struct window_relative { static constexpr bool value = false; };
template <class V>
struct window_relative<Win16Entries<V>> { static constexpr bool value = true; };
template <class V>
struct window_relative<Win16Records<V>> { static constexpr bool value = true; };

// rec[k] = off16[k] | value bits << 16 (the staged kernels' 4-byte records of 16-bit values).
template <class V>
__global__ void pack_w16(int64_t nnz, const uint16_t* __restrict__ off16, const V* __restrict__ val,
                         uint32_t* __restrict__ rec) {
    const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k < nnz) rec[k] = static_cast<uint32_t>(off16[k]) | (static_cast<uint32_t>(val[k].bits) << 16);
}

// Loads in flight per lane in one pass of the streamed SpMV loops: every lane issues all of
// its entry loads of a pass, then all of its gathers. Tiles hold a few entries per lane
// (about 3 per lane and column block on the 1M x 2M LP), so issuing them together, rather
// than one load-gather-store chain at a time, is what hides the gather latency: the FP32
// split A x_bar partials took 72.4 us with 8 per pass against 76.2 us with a 4-deep loop
// and a one-at-a-time remainder; 16 per pass (84 VGPRs, lower occupancy) took 74.2 us.
template <class T>
constexpr int stream_depth() { return sizeof(T) == 8 ? 4 : 8; }

// The fused tile-streamed steps (batches, and single LPs without a split) keep a 4-deep
// loop with a one-at-a-time remainder: their tiles hold 6-15 entries per lane and, in the
// batches, gather from vectors past the L2 capacity, and there the predicated passes were
// slower (512 random LPs: 19.87 s against 19.31 s, the same iterations).
template <class Ent, class T>
__device__ __forceinline__ void stream_products_unrolled(int s0, int s1, const Ent& ent, const T* __restrict__ vec,
                                                         T* __restrict__ lds) {
    constexpr int U = 4;
    int p = s0 + static_cast<int>(threadIdx.x);
    for (; p + (U - 1) * kBlock < s1; p += U * kBlock) {
        int c[U];
        T a[U];
#pragma unroll
        for (int u = 0; u < U; ++u) ent.load(p + u * kBlock, c[u], a[u]);
#pragma unroll
        for (int u = 0; u < U; ++u) lds[p + u * kBlock - s0] = a[u] * vec[c[u]];
    }
    for (; p < s1; p += kBlock) {
        int c;
        T a;
        ent.load(p, c, a);
        lds[p - s0] = a * vec[c];
    }
}

template <class Ent, class T>
__device__ __forceinline__ void stream_products(int s0, int s1, const Ent& ent, const T* __restrict__ vec,
                                                T* __restrict__ lds) {
    constexpr int U = stream_depth<T>();
    for (int p = s0 + static_cast<int>(threadIdx.x); p < s1; p += U * kBlock) {
        int c[U];
        T a[U], g[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            c[u] = 0;
            a[u] = 0;
            if (p + u * kBlock < s1) ent.load(p + u * kBlock, c[u], a[u]);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) g[u] = p + u * kBlock < s1 ? vec[c[u]] : static_cast<T>(0);
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (p + u * kBlock < s1) lds[p + u * kBlock - s0] = a[u] * g[u];
    }
}

// Whole-workgroup sum of one long row (fixed order: strided lane sums, then a tree).
template <class Ent, class T>
__device__ __forceinline__ T long_row_sum(int s0, int s1, const Ent& ent, const T* __restrict__ vec,
                                          T* __restrict__ lds) {
    T acc = 0;
    for (int p = s0 + static_cast<int>(threadIdx.x); p < s1; p += kBlock) {
        int c;
        T a;
        ent.load(p, c, a);
        acc += a * vec[c];
    }
    lds[threadIdx.x] = acc;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if (static_cast<int>(threadIdx.x) < w) lds[threadIdx.x] += lds[threadIdx.x + w];
        __syncthreads();
    }
    const T r = lds[0];
    __syncthreads();
    return r;
}

template <class Ent, class T, bool MULTI>
__global__ __launch_bounds__(kBlock) void stream_primal_step(
    int n, const int* __restrict__ list, const int* __restrict__ tile_row, const int* __restrict__ tp, Ent ent,
    const T* __restrict__ y, const T* __restrict__ dhat, const T* __restrict__ lcs,
    const T* __restrict__ ucs, T* __restrict__ x, T* __restrict__ xbar, const T* __restrict__ ax, const double* __restrict__ xhat,
    double* __restrict__ xfull, const InstState* __restrict__ st, const int* __restrict__ col_inst,
    double* __restrict__ part, int it, int flags, double rho) {
    __shared__ T lds[kStreamCap];
    const int tile = list ? list[blockIdx.x] : static_cast<int>(blockIdx.x);   // batches: the live tiles
    const int r0 = tile_row[tile], r1 = tile_row[tile + 1];
    int f = 0, l = 0;
    if (MULTI) {
        f = col_inst[r0];
        l = col_inst[r1 - 1];
    }
    bool any = false;
    for (int q = f; q <= l; ++q) any |= (st[q].active != 0);
    if (!any) return;
    const int s0 = tp[r0], s1 = tp[r1];
    const int j = r0 + static_cast<int>(threadIdx.x);
    const int my = (MULTI && j < r1) ? col_inst[j] : 0;
    double c[3] = {0.0, 0.0, 0.0};
    if (s1 - s0 > kStreamCap) {  // one long row
        const T acc = long_row_sum<Ent, T>(s0, s1, ent, y, lds);
        if (threadIdx.x == 0 && (!MULTI || st[my].active != 0))
            primal_epilogue<T>(acc, r0, st[my], it, flags, rho, dhat, lcs, ucs, x, xbar, ax, xhat, xfull, c);
    } else {
        // the row's operands and segment bounds are loaded before the gathers
        const bool mine = j < r1 && (!MULTI || st[my].active != 0);
        PrimalOps<T> o{};
        int q0 = 0, q1 = 0;
        if (mine) {
            o = primal_ops<T>(j, x, xbar, ax, dhat, lcs, ucs, st[my], it, rho);
            q0 = tp[j] - s0;
            q1 = tp[j + 1] - s0;
        }
        stream_products_unrolled<Ent, T>(s0, s1, ent, y, lds);
        __syncthreads();
        if (mine) {
            T acc = 0;
            for (int q = q0; q < q1; ++q) acc += lds[q];
            primal_update<T>(acc, o, j, st[my], it, flags, rho, x, xbar, ax, xhat, xfull, c);
        }
    }
    if (flags) {
        const int mask = ((flags & 1) ? 1 : 0) | ((flags & 2) ? 6 : 0);
        write_tile_partials<3, MULTI>(c, my, f, l, tile, part, mask);
    }
}

template <class Ent, class T, bool MULTI>
__global__ __launch_bounds__(kBlock) void stream_dual_step(
    int m, const int* __restrict__ list, const int* __restrict__ tile_row, const int* __restrict__ rp, Ent ent,
    const T* __restrict__ xbar, const T* __restrict__ yhat_t, const T* __restrict__ lrs,
    const T* __restrict__ urs, T* __restrict__ y, const T* __restrict__ ay, const double* __restrict__ yhat, double* __restrict__ yfull,
    const InstState* __restrict__ st, const int* __restrict__ row_inst, double* __restrict__ part, int it, int flags,
    double rho) {
    __shared__ T lds[kStreamCap];
    const int tile = list ? list[blockIdx.x] : static_cast<int>(blockIdx.x);
    const int r0 = tile_row[tile], r1 = tile_row[tile + 1];
    int f = 0, l = 0;
    if (MULTI) {
        f = row_inst[r0];
        l = row_inst[r1 - 1];
    }
    bool any = false;
    for (int q = f; q <= l; ++q) any |= (st[q].active != 0);
    if (!any) return;
    const int s0 = rp[r0], s1 = rp[r1];
    const int i = r0 + static_cast<int>(threadIdx.x);
    const int my = (MULTI && i < r1) ? row_inst[i] : 0;
    double c[3] = {0.0, 0.0, 0.0};
    if (s1 - s0 > kStreamCap) {
        const T w = long_row_sum<Ent, T>(s0, s1, ent, xbar, lds);
        if (threadIdx.x == 0 && (!MULTI || st[my].active != 0))
            dual_epilogue<T>(w, r0, st[my], it, flags, rho, yhat_t, lrs, urs, y, ay, yhat, yfull, c);
    } else {
        // the row's operands and segment bounds are loaded before the gathers
        const bool mine = i < r1 && (!MULTI || st[my].active != 0);
        DualOps<T> o{};
        int q0 = 0, q1 = 0;
        if (mine) {
            o = dual_ops<T>(i, y, yhat_t, lrs, urs);
            q0 = rp[i] - s0;
            q1 = rp[i + 1] - s0;
        }
        stream_products_unrolled<Ent, T>(s0, s1, ent, xbar, lds);
        __syncthreads();
        if (mine) {
            T w = 0;
            for (int q = q0; q < q1; ++q) w += lds[q];
            dual_update<T>(w, o, i, st[my], it, flags, rho, y, ay, yhat, yfull, c);
        }
    }
    if (flags) {
        const int mask = ((flags & 1) ? 1 : 0) | ((flags & 2) ? 6 : 0);
        write_tile_partials<3, MULTI>(c, my, f, l, tile, part, mask);
    }
}

// ---------------------------------------------------------------------------------
// LDS-staged fused steps, for batches of small LPs. A batched tile-streamed step pays one L2
// line per nonzero for its gathers, and that line traffic, not HBM, sets its speed. Here every
// tile holds rows of a single instance (at most RPL * NT rows and stage_cap<T, NT>() entries,
// built on the device) and carries the window [lo, hi) of columns its entries touch. The
// workgroup copies the window of the gathered vector into LDS with coalesced loads (chunk by
// chunk when it is wider than the buffer) and each entry then reads its vector element from
// LDS. A lane keeps its stage_depth<T>() entries in registers across the chunks; every entry
// is multiplied in exactly one chunk. The products then go to the same buffer by entry
// position and one lane per row adds its segment in index order: the row sums are those of
// the tile-streamed kernels, bit for bit. Workgroups of NT = 1024 or 256 threads (the
// autotuner times both: the 256-thread shape keeps eight workgroups per CU, which paid on the
// staircase batch, whose windows are narrow; the 1024-thread shape's larger tiles paid on the
// random batch) hold NT * 64 bytes of LDS, 64 KB or 16 KB, which leaves room for 32 waves per
// CU; the kernels keep them by asking for 8 waves per SIMD (at most 64 VGPRs: the first build,
// with predicated entry loads, spilled to scratch and ran 9x slower). A lane holds 64 bytes of
// entries (16 FP32 or 8 FP64) in either shape.
// Workgroup b runs tile list[b]: the host launches the tiles of the instances still active
// (the list shrinks as instances finish; each workgroup still checks its instance's state),
// because only 512 (2,048) workgroups are resident and workgroups of finished instances, each
// exiting after two loads, cost about 64 us per iteration in the tail of a 512-LP batch.
constexpr int kStageLaneBytes = 64;
template <class T, int NT>
constexpr int stage_cap() { return NT * kStageLaneBytes / static_cast<int>(sizeof(T)); }
template <class T>
constexpr int stage_depth() { return kStageLaneBytes / static_cast<int>(sizeof(T)); }

// lds[q] = vec[base + q] for q < len, with direct global-to-LDS loads (global_load_lds_dword):
// wave w copies 4-byte words q0 .. q0+63 (q0 = 64 w, 64 w + NT, ...), lane l word q0 + l, into
// LDS at the wave-uniform base lds + q0. The loads occupy no VGPRs, so every word of a lane is
// in flight at once (a copy through registers, four loads at a time, waited for each group);
// the caller's barrier waits for them (vmcnt). Lanes past the end re-read the last word and
// write it to LDS words past len, which no entry reads (the buffer holds whole waves' words).
template <class T, int NT>
__device__ __forceinline__ void stage_window(const T* __restrict__ vec, int base, int len, T* __restrict__ lds) {
    constexpr int W = static_cast<int>(sizeof(T) / 4);
    const uint32_t* src = reinterpret_cast<const uint32_t*>(vec) + static_cast<int64_t>(base) * W;
    uint32_t* dst = reinterpret_cast<uint32_t*>(lds);
    const int words = len * W;
    const int lane = static_cast<int>(threadIdx.x) & (kWave - 1);
    for (int q0 = static_cast<int>(threadIdx.x) - lane; q0 < words; q0 += NT) {
        const int q = q0 + lane < words ? q0 + lane : words - 1;
        __builtin_amdgcn_global_load_lds(src + q, dst + q0, 4, 0, 0);
    }
}

// Products of the tile's entries [s0, s1) (at most stage_cap<T, NT>()) with vec, whose elements
// the entries read in the window [wlo, whi); on return lds[p - s0] = a_p * vec[col_p]. Entry
// columns are absolute, or relative to wlo for the 16-bit offset format.
template <class Ent, class T, int NT>
__device__ __forceinline__ void staged_products(int s0, int s1, int wlo, int whi, const Ent& ent,
                                                const T* __restrict__ vec, T* __restrict__ lds) {
    constexpr int U = stage_depth<T>();
    constexpr int CAP = stage_cap<T, NT>();
    if (s1 <= s0) return;   // no entries (uniform over the workgroup)
    int c[U];
    T a[U];
    // every lane loads all U slots from an in-bounds position (the tile's last entry for the
    // slots past its end) and marks the extra ones with column -1, whose offset from any chunk
    // base is out of range: unconditional loads keep the slots in plain registers
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const int p = s0 + static_cast<int>(threadIdx.x) + u * NT;
        int cc;
        T aa;
        ent.load(p < s1 ? p : s1 - 1, cc, aa);
        c[u] = p < s1 ? cc : -1;
        a[u] = aa;
    }
    for (int base = wlo; base < whi; base += CAP) {
        const int len = whi - base < CAP ? whi - base : CAP;
        if (base != wlo) __syncthreads();   // the previous chunk has been read
        stage_window<T, NT>(vec, base, len, lds);
        __syncthreads();
        const int cb = window_relative<Ent>::value ? base - wlo : base;
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const unsigned d = static_cast<unsigned>(c[u] - cb);
            if (d < static_cast<unsigned>(len)) a[u] = a[u] * lds[d];
        }
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const int p = s0 + static_cast<int>(threadIdx.x) + u * NT;
        if (p < s1) lds[p - s0] = a[u];
    }
    __syncthreads();
}

// Whole-workgroup sum of one row longer than the buffer (gathers from memory; fixed order).
template <class Ent, class T, int NT>
__device__ __forceinline__ T staged_long_row(int s0, int s1, const Ent& ent, const T* __restrict__ vec,
                                             T* __restrict__ lds) {
    T acc = 0;
    for (int p = s0 + static_cast<int>(threadIdx.x); p < s1; p += NT) {
        int c;
        T a;
        ent.load(p, c, a);
        acc += a * vec[c];
    }
    lds[threadIdx.x] = acc;
    __syncthreads();
    for (int w = NT / 2; w > 0; w >>= 1) {
        if (static_cast<int>(threadIdx.x) < w) lds[threadIdx.x] += lds[threadIdx.x + w];
        __syncthreads();
    }
    const T r = lds[0];
    __syncthreads();
    return r;
}

// The tile's residual partials (one instance per tile) to part[slot * 3 + q].
template <int NT>
__device__ __forceinline__ void stage_partials(const double (&v)[3], int slot, double* __restrict__ part, int mask,
                                               double (*red)[3]) {
    double t[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) t[q] = group_sum<kWave>(v[q]);
    const int wave = threadIdx.x / kWave;
    if ((threadIdx.x & (kWave - 1)) == 0) {
#pragma unroll
        for (int q = 0; q < 3; ++q) red[wave][q] = t[q];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            double s = red[0][q];
            for (int w = 1; w < NT / kWave; ++w) s += red[w][q];
            if ((mask >> q) & 1) part[static_cast<int64_t>(slot) * 3 + q] = s;
        }
    }
}

// RPL rows per lane: tiles of up to RPL * NT rows. Matrices with short rows (the columns of a
// random batch instance have 6 entries) take RPL = 2, so that a tile still holds enough
// entries to pay for copying its window.
template <class Ent, class T, bool MULTI, int RPL, int NT>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(8))) void staged_primal_step(
    const int* __restrict__ list, const int4* __restrict__ desc, const int* __restrict__ tp, Ent ent, const T* __restrict__ y, const T* __restrict__ dhat,
    const T* __restrict__ lcs, const T* __restrict__ ucs, T* __restrict__ x, T* __restrict__ xbar, const T* __restrict__ ax,
    const double* __restrict__ xhat, double* __restrict__ xfull, const InstState* __restrict__ st,
    double* __restrict__ part, int it, int flags, double rho) {
    __shared__ T lds[stage_cap<T, NT>()];
    __shared__ double red[NT / kWave][3];
    const int tile = list[blockIdx.x];
    const int4 d0 = desc[2 * tile], d1 = desc[2 * tile + 1];   // rows, entries | window, instance
    const int my = MULTI ? d1.z : 0;   // every row of the tile belongs to this instance
    if (st[my].active == 0) return;
    const int r0 = d0.x, r1 = d0.y, s0 = d0.z, s1 = d0.w;
    double c[3] = {0.0, 0.0, 0.0};
    if (s1 - s0 > stage_cap<T, NT>()) {   // one long row
        const T acc = staged_long_row<Ent, T, NT>(s0, s1, ent, y, lds);
        if (threadIdx.x == 0) primal_epilogue<T>(acc, r0, st[my], it, flags, rho, dhat, lcs, ucs, x, xbar, ax, xhat, xfull, c);
    } else {
        // the rows' operands and bounds are loaded before the staging (a lane past the tile's end
        // loads its last row's and skips the update)
        PrimalOps<T> o[RPL];
        int q0[RPL], q1[RPL];
#pragma unroll
        for (int u = 0; u < RPL; ++u) {
            const int j = r0 + static_cast<int>(threadIdx.x) + u * NT;
            const int jc = j < r1 ? j : r1 - 1;
            o[u] = primal_ops<T>(jc, x, xbar, ax, dhat, lcs, ucs, st[my], it, rho);
            q0[u] = tp[jc] - s0;
            q1[u] = tp[jc + 1] - s0;
        }
        staged_products<Ent, T, NT>(s0, s1, d1.x, d1.y, ent, y, lds);
#pragma unroll
        for (int u = 0; u < RPL; ++u) {
            const int j = r0 + static_cast<int>(threadIdx.x) + u * NT;
            if (j < r1) {
                T acc = 0;
                for (int q = q0[u]; q < q1[u]; ++q) acc += lds[q];
                primal_update<T>(acc, o[u], j, st[my], it, flags, rho, x, xbar, ax, xhat, xfull, c);
            }
        }
    }
    if (flags) {
        const int mask = ((flags & 1) ? 1 : 0) | ((flags & 2) ? 6 : 0);
        stage_partials<NT>(c, tile + my, part, mask, red);
    }
}

template <class Ent, class T, bool MULTI, int RPL, int NT>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(8))) void staged_dual_step(
    const int* __restrict__ list, const int4* __restrict__ desc, const int* __restrict__ rp, Ent ent, const T* __restrict__ xbar, const T* __restrict__ yhat_t,
    const T* __restrict__ lrs, const T* __restrict__ urs, T* __restrict__ y, const T* __restrict__ ay, const double* __restrict__ yhat,
    double* __restrict__ yfull, const InstState* __restrict__ st, double* __restrict__ part, int it, int flags,
    double rho) {
    __shared__ T lds[stage_cap<T, NT>()];
    __shared__ double red[NT / kWave][3];
    const int tile = list[blockIdx.x];
    const int4 d0 = desc[2 * tile], d1 = desc[2 * tile + 1];
    const int my = MULTI ? d1.z : 0;
    if (st[my].active == 0) return;
    const int r0 = d0.x, r1 = d0.y, s0 = d0.z, s1 = d0.w;
    double c[3] = {0.0, 0.0, 0.0};
    if (s1 - s0 > stage_cap<T, NT>()) {
        const T w = staged_long_row<Ent, T, NT>(s0, s1, ent, xbar, lds);
        if (threadIdx.x == 0) dual_epilogue<T>(w, r0, st[my], it, flags, rho, yhat_t, lrs, urs, y, ay, yhat, yfull, c);
    } else {
        DualOps<T> o[RPL];
        int q0[RPL], q1[RPL];
#pragma unroll
        for (int u = 0; u < RPL; ++u) {
            const int i = r0 + static_cast<int>(threadIdx.x) + u * NT;
            const int ic = i < r1 ? i : r1 - 1;
            o[u] = dual_ops<T>(ic, y, yhat_t, lrs, urs);
            q0[u] = rp[ic] - s0;
            q1[u] = rp[ic + 1] - s0;
        }
        staged_products<Ent, T, NT>(s0, s1, d1.x, d1.y, ent, xbar, lds);
#pragma unroll
        for (int u = 0; u < RPL; ++u) {
            const int i = r0 + static_cast<int>(threadIdx.x) + u * NT;
            if (i < r1) {
                T w = 0;
                for (int q = q0[u]; q < q1[u]; ++q) w += lds[q];
                dual_update<T>(w, o[u], i, st[my], it, flags, rho, y, ay, yhat, yfull, c);
            }
        }
    }
    if (flags) {
        const int mask = ((flags & 1) ? 1 : 0) | ((flags & 2) ? 6 : 0);
        stage_partials<NT>(c, tile + my, part, mask, red);
    }
}

// ---------------------------------------------------------------------------------
// Row stride of the S partial-sum vectors of a column split: a multiple of 64 entries, so that
// every vector starts 256-byte aligned and the finish kernels read it 16 bytes at a time.
__host__ __device__ constexpr int64_t split_pstride(int rows) { return (static_cast<int64_t>(rows) + 63) / 64 * 64; }

// XCD-aware column-split SpMV (single LP). The matrix is stored as S column blocks,
// each a CSR over all rows with absolute pointers bp[k*(rows+1) + r]. Workgroup w works on
// block w % S: consecutive workgroup ids are dispatched round-robin over the 8 XCDs, so
// each XCD (with S dividing 8) only ever gathers from one block of the vector, which stays
// resident in that XCD's 4 MB L2. Block partial sums go to part[k*pstride + r]; the finish
// kernels add the S partials in block order (deterministic) and run the epilogue.
template <class V, class T, int G>
__global__ __launch_bounds__(kBlock) void split_partials(int rows, int S, const int* __restrict__ bp,
                                                         const int* __restrict__ bidx, const V* __restrict__ bval,
                                                         const T* __restrict__ vec, T* __restrict__ part,
                                                         const InstState* __restrict__ st) {
    if (st[0].active == 0) return;
    constexpr int R = kBlock / G;
    const int k = static_cast<int>(blockIdx.x) % S;
    const int t = static_cast<int>(blockIdx.x) / S;
    const int r = t * R + static_cast<int>(threadIdx.x) / G;
    const int lane = threadIdx.x % G;
    T acc = 0;
    if (r < rows) {
        const int* p = bp + static_cast<int64_t>(k) * (rows + 1);
        const int q1 = p[r + 1];
        for (int q = p[r] + lane; q < q1; q += G) acc += static_cast<T>(load_val(bval, q)) * vec[load_idx(bidx, q)];
    }
    acc = group_sum<G>(acc);
    if (r < rows && lane == 0) part[static_cast<int64_t>(k) * split_pstride(rows) + r] = acc;
}

// Tile-streamed column-split partials. Workgroup w handles block k = w % S (the same XCD
// mapping as split_partials) and row tile t = w / S. The tiles (built on the device) are common
// to all blocks: at most split_rows<T>() rows, and at most split_cap<T>() entries in every block; a
// single row longer than that in some block is a tile of its own, reduced by the whole
// workgroup.
// The block's span of values and indices is streamed with coalesced loads independent of
// the row structure, the products go to LDS, and each thread adds up to split_rows<T>()/256
// rows of the tile in index order (deterministic).
// Rows per tile: 256 for 4-byte iterates, 512 for FP64. Shorter tiles balance the CUs better;
// on the 1M x 2M LP (about 3 entries per row and block) the FP32 A x_bar step took 89.2 us at
// 128 rows, 82.8 at 256, 84.2 at 512, 87.0 at 1024 and 92 at 2048, the FP64 steps 119.1 + 112.8 us at 512
// rows, 120.4 + 114.3 at 256 and 122.1 + 115.6 at 1024.
template <class T>
constexpr int split_rows() { return sizeof(T) == 8 ? 512 : 256; }
// LDS product buffer of one split tile: 16 KB (4096 FP32 / 2048 FP64 entries), so that
// eight workgroups of four waves fit a CU in either precision (8 and 32 KB measured the same).
constexpr int kSplitLdsBytes = 16384;
template <class T>
constexpr int split_cap() { return kSplitLdsBytes / static_cast<int>(sizeof(T)); }

// tile_base[t * S + k] = first entry of tile t in block k (t = 0..tiles): the tile's span of
// block k comes from one load that depends on nothing, so the entry stream starts without
// waiting for the tile's rows (70.9 us against 72.4 us for the FP32 A x_bar partials when
// the tile's span came from the block pointers of its first and last rows).
// toff[k * rows + r] = first entry of row r in block k relative to its tile's first entry
// (at most split_cap<T>() in a tile that fits LDS; a long-row tile's single row has 0): the
// row bounds cost two bytes per row and block instead of a 4-byte block pointer. They are
// loaded after the stream is issued: loaded before it, they made the stream wait for the
// tile's row range (83.6 us against 77.9 us for the FP32 A x_bar pair).
template <class Ent, class T>
__global__ __launch_bounds__(kBlock) void split_stream_partials(int rows, int S, const int* __restrict__ tile_row,
                                                                const int* __restrict__ tile_base,
                                                                const uint16_t* __restrict__ toff, Ent ent,
                                                                const T* __restrict__ vec, T* __restrict__ part,
                                                                const InstState* __restrict__ st) {
    __shared__ T lds[split_cap<T>()];
    if (st[0].active == 0) return;
    const int k = static_cast<int>(blockIdx.x) % S;
    const int t = static_cast<int>(blockIdx.x) / S;
    const int s0 = tile_base[t * S + k], s1 = tile_base[(t + 1) * S + k];
    const int r0 = tile_row[t], r1 = tile_row[t + 1];
    T* out = part + static_cast<int64_t>(k) * split_pstride(rows);
    if (s1 - s0 > split_cap<T>()) {
        const T acc = long_row_sum<Ent, T>(s0, s1, ent, vec, lds);
        if (threadIdx.x == 0) out[r0] = acc;
        return;
    }
    stream_products<Ent, T>(s0, s1, ent, vec, lds);
    const uint16_t* o = toff + static_cast<int64_t>(k) * rows;
    constexpr int RPT = split_rows<T>() / kBlock;
    int lo[RPT], hi[RPT];
#pragma unroll
    for (int u = 0; u < RPT; ++u) {
        const int r = r0 + static_cast<int>(threadIdx.x) + u * kBlock;
        lo[u] = r < r1 ? static_cast<int>(o[r]) : 0;
        hi[u] = r + 1 < r1 ? static_cast<int>(o[r + 1]) : (r < r1 ? s1 - s0 : 0);
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < RPT; ++u) {
        const int r = r0 + static_cast<int>(threadIdx.x) + u * kBlock;
        if (r < r1) {
            T acc = 0;
            for (int q = lo[u]; q < hi[u]; ++q) acc += lds[q];
            out[r] = acc;
        }
    }
}

// The finish kernels add the S block partials of finish_rows<T>() consecutive rows per thread
// with 16-byte loads (one per partial vector and operand) and run the epilogue row by row,
// with the same arithmetic as the one-row kernels; a ragged end is done row by row.
template <class T>
constexpr int finish_rows() { return static_cast<int>(16 / sizeof(T)); }
template <class T>
struct Vec16;
template <>
struct Vec16<float> {
    typedef float type __attribute__((ext_vector_type(4)));
};
template <>
struct Vec16<double> {
    typedef double type __attribute__((ext_vector_type(2)));
};
template <class T>
__device__ __forceinline__ typename Vec16<T>::type load16(const T* p) {
    return *reinterpret_cast<const typename Vec16<T>::type*>(p);
}
template <class T>
__device__ __forceinline__ void store16(T* p, const typename Vec16<T>::type& v) {
    *reinterpret_cast<typename Vec16<T>::type*>(p) = v;
}

template <class T>
__global__ __launch_bounds__(kBlock) void primal_finish(
    int n, int S, const T* __restrict__ psum, const T* __restrict__ dhat, const T* __restrict__ lcs,
    const T* __restrict__ ucs, T* __restrict__ x, T* __restrict__ xbar, const T* __restrict__ ax, const double* __restrict__ xhat,
    double* __restrict__ xfull, const InstState* __restrict__ st, double* __restrict__ part, int it, int flags,
    double rho) {
    if (st[0].active == 0) return;
    constexpr int V = finish_rows<T>();
    typedef typename Vec16<T>::type VT;
    const int64_t ps = split_pstride(n);
    const int j0 = (static_cast<int>(blockIdx.x) * kBlock + static_cast<int>(threadIdx.x)) * V;
    const InstState& s = st[0];
    double c[3] = {0.0, 0.0, 0.0};
    if (j0 + V <= n) {
        VT acc = load16<T>(psum + j0);
        for (int k = 1; k < S; ++k) acc += load16<T>(psum + k * ps + j0);
        const VT xb = load16<T>(xbar + j0), dh = load16<T>(dhat + j0), lo = load16<T>(lcs + j0),
                 hi = load16<T>(ucs + j0);
        VT an = {};
        if (s.anchored) an = load16<T>(ax + j0);
        VT xk;
        if (rho == 1.0) {
            const T lam_prev = anchor_weight<T>(s, it - 1);   // x recovered from x_bar, as x_start
            if (s.anchored) {
                const T pull_prev = anchor_pull<T>(s, it - 1);
                for (int v = 0; v < V; ++v) xk[v] = lam_prev * xb[v] + pull_prev * an[v];
            } else {
                for (int v = 0; v < V; ++v) xk[v] = lam_prev * xb[v];
            }
        } else {
            xk = load16<T>(x + j0);
        }
        const T tau = static_cast<T>(s.tau), lam = anchor_weight<T>(s, it), pull = anchor_pull<T>(s, it);
        VT xb_new, x_new;
        for (int v = 0; v < V; ++v) {
            const T xt = primal_point<T>(acc[v], PrimalOps<T>{xk[v], dh[v], lo[v], hi[v]}, tau);
            xb_new[v] = static_cast<T>(2) * xt - xk[v];
            x_new[v] = lam * ((static_cast<T>(1) + static_cast<T>(rho)) * xt - static_cast<T>(rho) * xk[v]);
            if (s.anchored) x_new[v] += pull * an[v];
            primal_residuals<T>(xt, xk[v], an[v], j0 + v, flags, xhat, xfull, c);
        }
        store16<T>(xbar + j0, xb_new);
        if (rho != 1.0) store16<T>(x + j0, x_new);
    } else {
        for (int j = j0; j < n; ++j) {
            T acc = psum[j];
            for (int k = 1; k < S; ++k) acc += psum[k * ps + j];
            primal_epilogue<T>(acc, j, s, it, flags, rho, dhat, lcs, ucs, x, xbar, ax, xhat, xfull, c);
        }
    }
    if (flags) {
        const int mask = ((flags & 1) ? 1 : 0) | ((flags & 2) ? 6 : 0);
        write_tile_partials<3, false>(c, 0, 0, 0, blockIdx.x, part, mask);
    }
}

template <class T>
__global__ __launch_bounds__(kBlock) void dual_finish(
    int m, int S, const T* __restrict__ psum, const T* __restrict__ yhat_t, const T* __restrict__ lrs,
    const T* __restrict__ urs, T* __restrict__ y, const T* __restrict__ ay, const double* __restrict__ yhat, double* __restrict__ yfull,
    const InstState* __restrict__ st, double* __restrict__ part, int it, int flags, double rho) {
    if (st[0].active == 0) return;
    constexpr int V = finish_rows<T>();
    typedef typename Vec16<T>::type VT;
    const int64_t ps = split_pstride(m);
    const int i0 = (static_cast<int>(blockIdx.x) * kBlock + static_cast<int>(threadIdx.x)) * V;
    const InstState& s = st[0];
    double c[3] = {0.0, 0.0, 0.0};
    if (i0 + V <= m) {
        VT w = load16<T>(psum + i0);
        for (int k = 1; k < S; ++k) w += load16<T>(psum + k * ps + i0);
        const bool eq = lrs == urs;
        const VT yk = load16<T>(y + i0), lo = load16<T>(lrs + i0);
        VT hi = lo, yh = lo;
        if (!eq) {
            hi = load16<T>(urs + i0);
            yh = load16<T>(yhat_t + i0);
        }
        VT an = {};
        if (s.anchored) an = load16<T>(ay + i0);
        const T sig = static_cast<T>(s.sigma), lam = anchor_weight<T>(s, it), pull = anchor_pull<T>(s, it);
        const bool anchored = s.anchored != 0;
        VT y_new;
        for (int v = 0; v < V; ++v) {
            const T yt = dual_point<T>(w[v], DualOps<T>{yk[v], eq ? static_cast<T>(0) : yh[v], lo[v], hi[v], eq},
                                       sig);
            y_new[v] = dual_next<T>(yt, yk[v], an[v], lam, pull, anchored, rho);
            dual_residuals<T>(yt, yk[v], an[v], i0 + v, flags, yhat, yfull, c);
        }
        store16<T>(y + i0, y_new);
    } else {
        for (int i = i0; i < m; ++i) {
            T w = psum[i];
            for (int k = 1; k < S; ++k) w += psum[k * ps + i];
            dual_epilogue<T>(w, i, s, it, flags, rho, yhat_t, lrs, urs, y, ay, yhat, yfull, c);
        }
    }
    if (flags) {
        const int mask = ((flags & 1) ? 1 : 0) | ((flags & 2) ? 6 : 0);
        write_tile_partials<3, false>(c, 0, 0, 0, blockIdx.x, part, mask);
    }
}

// Gather of the scaled matrix into split or sliced-ELL order: out[q] = in[perm[q]] (0 where
// perm[q] < 0), converted to V.
template <class V>
__global__ void gather_split_vals(int64_t nnz, const int* __restrict__ perm, const double* __restrict__ in,
                                  V* __restrict__ out) {
    const int64_t q = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (q >= nnz) return;
    const int src = perm[q];
    const double v = src >= 0 ? in[src] : 0.0;   // a negative entry marks sliced-ELL padding
    if constexpr (sizeof(V) == 8) {
        out[q] = v;
    } else if constexpr (sizeof(V) == 4) {
        out[q] = static_cast<float>(v);
    } else if constexpr (std::is_same<V, f16_t>::value) {
        out[q].bits = f16_bits_from(v);
    } else {
        uint32_t u = __float_as_uint(static_cast<float>(v));
        if ((u & 0x7f800000u) != 0x7f800000u) u += 0x7fffu + ((u >> 16) & 1u);
        out[q].bits = static_cast<uint16_t>(u >> 16);
    }
}

// ---------------------------------------------------------------------------------
// FP64 KKT passes (the check, and the audit on original data when the unscale arrays
// are null). Row pass: ax = A x; q0 = sum r_p^2 (original units), q1 = row part of the
// dual objective, q2 = row-dual sign violations^2 (original units).
// ---------------------------------------------------------------------------------
// Sliced-ELL fused steps (single LP). Rows come in slices of 64, one per wavefront; a slice
// stores its rows' entries column-major, padded to its longest row with zero-valued entries
// that point at the row's first column. Lane r of the wavefront reads entry q of its row at
// off[s] + 64 q + r: one coalesced load per q, no row pointers, no LDS. The row sum runs in
// index order. The autotuner offers it when padding adds at most a quarter to the entries,
// as when every row has the same length.

template <class Ent, class T>
__device__ __forceinline__ T ell_row_sum(int r, const int* __restrict__ off, const int* __restrict__ width,
                                         const Ent& ent, const T* __restrict__ vec) {
    const int sl = r / kEllSlice;
    const int64_t base = static_cast<int64_t>(off[sl]) + (r % kEllSlice);
    const int w = width[sl];
    // passes of up to stream_depth<T>() entries: all loads of a pass, then all its gathers
    // (the FP32 ELL primal step of the 1M x 2M LP, 6 entries per row: 77-79 us against 82 us
    // for a 4-deep loop with a one-at-a-time remainder)
    constexpr int U = stream_depth<T>();
    T acc = 0;
    for (int q0 = 0; q0 < w; q0 += U) {
        int c[U];
        T a[U], g[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            c[u] = 0;
            a[u] = 0;
            if (q0 + u < w) ent.load(base + static_cast<int64_t>(q0 + u) * kEllSlice, c[u], a[u]);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) g[u] = q0 + u < w ? vec[c[u]] : static_cast<T>(0);
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (q0 + u < w) acc += a[u] * g[u];
    }
    return acc;
}

template <class Ent, class T>
__global__ __launch_bounds__(kBlock) void ell_primal_step(
    int n, const int* __restrict__ off, const int* __restrict__ width, Ent ent, const T* __restrict__ y,
    const T* __restrict__ dhat, const T* __restrict__ lcs, const T* __restrict__ ucs, T* __restrict__ x,
    T* __restrict__ xbar, const T* __restrict__ ax, const double* __restrict__ xhat, double* __restrict__ xfull,
    const InstState* __restrict__ st, double* __restrict__ part, int it, int flags, double rho) {
    if (st[0].active == 0) return;
    const int j = static_cast<int>(blockIdx.x) * kBlock + static_cast<int>(threadIdx.x);
    double c[3] = {0.0, 0.0, 0.0};
    if (j < n) {
        const PrimalOps<T> o = primal_ops<T>(j, x, xbar, ax, dhat, lcs, ucs, st[0], it, rho);   // before the gathers
        const T acc = ell_row_sum<Ent, T>(j, off, width, ent, y);
        primal_update<T>(acc, o, j, st[0], it, flags, rho, x, xbar, ax, xhat, xfull, c);
    }
    if (flags) {
        const int mask = ((flags & 1) ? 1 : 0) | ((flags & 2) ? 6 : 0);
        write_tile_partials<3, false>(c, 0, 0, 0, blockIdx.x, part, mask);
    }
}

template <class Ent, class T>
__global__ __launch_bounds__(kBlock) void ell_dual_step(
    int m, const int* __restrict__ off, const int* __restrict__ width, Ent ent, const T* __restrict__ xbar,
    const T* __restrict__ yhat_t, const T* __restrict__ lrs, const T* __restrict__ urs, T* __restrict__ y, const T* __restrict__ ay,
    const double* __restrict__ yhat, double* __restrict__ yfull, const InstState* __restrict__ st,
    double* __restrict__ part, int it, int flags, double rho) {
    if (st[0].active == 0) return;
    const int i = static_cast<int>(blockIdx.x) * kBlock + static_cast<int>(threadIdx.x);
    double c[3] = {0.0, 0.0, 0.0};
    if (i < m) {
        const DualOps<T> o = dual_ops<T>(i, y, yhat_t, lrs, urs);
        const T w = ell_row_sum<Ent, T>(i, off, width, ent, xbar);
        dual_update<T>(w, o, i, st[0], it, flags, rho, y, ay, yhat, yfull, c);
    }
    if (flags) {
        const int mask = ((flags & 1) ? 1 : 0) | ((flags & 2) ? 6 : 0);
        write_tile_partials<3, false>(c, 0, 0, 0, blockIdx.x, part, mask);
    }
}

// Interleave an index array and an FP32 value array into 8-byte (index, value) records.
__global__ void pack_records(int64_t nnz, const int* __restrict__ idx, const float* __restrict__ val,
                             vj_i32x2* __restrict__ rec) {
    const int64_t q = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (q >= nnz) return;
    vj_i32x2 r;
    r.x = idx[q];
    r.y = __float_as_int(val[q]);
    rec[q] = r;
}

template <int G, bool MULTI>
__global__ __launch_bounds__(kBlock) void kkt_rows(
    int m, const int* __restrict__ rp, const int* __restrict__ ci, const double* __restrict__ av,
    const double* __restrict__ x, const double* __restrict__ y, const double* __restrict__ lr,
    const double* __restrict__ ur, const double* __restrict__ ir, const double* __restrict__ py,
    double* __restrict__ ax_out, const InstState* __restrict__ st, int only_active,
    const int* __restrict__ row_inst, double* __restrict__ part) {
    constexpr int R = kBlock / G;
    const int tile = blockIdx.x;
    const int r0 = tile * R;
    const int i = r0 + static_cast<int>(threadIdx.x) / G;
    const int lane = threadIdx.x % G;
    int f = 0, l = 0;
    if (MULTI) {
        f = row_inst[r0];
        l = row_inst[(r0 + R < m ? r0 + R : m) - 1];
    }
    if (only_active) {
        bool any = false;
        for (int q = f; q <= l; ++q) any |= (st[q].active != 0);
        if (!any) return;
    }
    const int my = (MULTI && i < m) ? row_inst[i] : 0;
    const bool act = (i < m) && (!only_active || st[my].active != 0);
    double acc = 0.0;
    if (act)
        for (int p = rp[i] + lane; p < rp[i + 1]; p += G) acc += av[p] * x[ci[p]];
    acc = group_sum<G>(acc);
    double c[3] = {0.0, 0.0, 0.0};
    if (act && lane == 0) {
        if (ax_out) ax_out[i] = acc;
        const double lo = lr[i], hi = ur[i];
        const bool lf = lo > -kInfBound, uf = hi < kInfBound;
        double r = 0.0;
        if (lf && acc < lo) r = acc - lo;
        if (uf && acc > hi) r = acc - hi;
        if (ir) r *= ir[i];
        c[0] = r * r;
        const double yi = y[i];
        double sv = 0.0;
        if (yi > 0.0) {
            if (lf) c[1] = yi * lo; else sv = yi;
        } else if (yi < 0.0) {
            if (uf) c[1] = yi * hi; else sv = yi;
        }
        if (py) sv *= py[i];
        c[2] = sv * sv;
    }
    write_tile_partials<3, MULTI>(c, my, f, l, tile, part, 7);
}

// Column pass: aty = A^T y; d = c - aty; lambda = projection of d onto the reduced-cost
// cone of the bounds; q0 = |d - lambda|^2 (original units), q1 = lambda . bounds,
// q2 = c . x.
template <int G, bool MULTI>
__global__ __launch_bounds__(kBlock) void kkt_cols(
    int n, const int* __restrict__ tp, const int* __restrict__ ti, const double* __restrict__ tv,
    const double* __restrict__ y, const double* __restrict__ x, const double* __restrict__ c,
    const double* __restrict__ lc, const double* __restrict__ uc, const double* __restrict__ ic,
    double* __restrict__ aty_out, const InstState* __restrict__ st, int only_active,
    const int* __restrict__ col_inst, double* __restrict__ part) {
    constexpr int R = kBlock / G;
    const int tile = blockIdx.x;
    const int r0 = tile * R;
    const int j = r0 + static_cast<int>(threadIdx.x) / G;
    const int lane = threadIdx.x % G;
    int f = 0, l = 0;
    if (MULTI) {
        f = col_inst[r0];
        l = col_inst[(r0 + R < n ? r0 + R : n) - 1];
    }
    if (only_active) {
        bool any = false;
        for (int q = f; q <= l; ++q) any |= (st[q].active != 0);
        if (!any) return;
    }
    const int my = (MULTI && j < n) ? col_inst[j] : 0;
    const bool act = (j < n) && (!only_active || st[my].active != 0);
    double acc = 0.0;
    if (act)
        for (int p = tp[j] + lane; p < tp[j + 1]; p += G) acc += tv[p] * y[ti[p]];
    acc = group_sum<G>(acc);
    double cc[3] = {0.0, 0.0, 0.0};
    if (act && lane == 0) {
        if (aty_out) aty_out[j] = acc;
        const double d = c[j] - acc;
        const double lo = lc[j], hi = uc[j];
        double lam = 0.0;
        if (d > 0.0 && lo > -kInfBound) lam = d;
        else if (d < 0.0 && hi < kInfBound) lam = d;
        double rd = d - lam;
        if (ic) rd *= ic[j];
        cc[0] = rd * rd;
        if (lam > 0.0) cc[1] = lam * lo;
        else if (lam < 0.0) cc[1] = lam * hi;
        cc[2] = c[j] * x[j];
    }
    write_tile_partials<3, MULTI>(cc, my, f, l, tile, part, 7);
}

// Per-instance rel-KKT from the two passes (audit: factor = 1, no state update).
// out8 per instance: rel_kkt, rel_p, rel_d, rel_gap, pobj, dobj, |r_p|, |r_d|.
__global__ __launch_bounds__(kBlock) void kkt_finish(
    const double* __restrict__ prow, int Rr, const double* __restrict__ pcol, int Rc,
    const int* __restrict__ row_off, const int* __restrict__ col_off,
    const double* __restrict__ norm_b, const double* __restrict__ norm_c,
    const double* __restrict__ offset, const double* __restrict__ objfac, double* __restrict__ out8) {
    __shared__ double lds[kBlock];
    const int inst = blockIdx.x;
    const int64_t r_a = row_off[inst], r_b = row_off[inst + 1];
    const int64_t c_a = col_off[inst], c_b = col_off[inst + 1];
    const double rp2 = reduce_instance(prow, 3, 0, inst, r_a, r_b, Rr, lds);
    const double dro = reduce_instance(prow, 3, 1, inst, r_a, r_b, Rr, lds);
    const double rsv = reduce_instance(prow, 3, 2, inst, r_a, r_b, Rr, lds);
    const double rdc = reduce_instance(pcol, 3, 0, inst, c_a, c_b, Rc, lds);
    const double dco = reduce_instance(pcol, 3, 1, inst, c_a, c_b, Rc, lds);
    const double po = reduce_instance(pcol, 3, 2, inst, c_a, c_b, Rc, lds);
    if (threadIdx.x == 0) {
        const double fac = objfac ? objfac[inst] : 1.0;
        const double pobj = offset[inst] + po * fac;
        const double dobj = offset[inst] + (dro + dco) * fac;
        const double nrp = sqrt(rp2), nrd = sqrt(rdc + rsv);
        const double rel_p = nrp / (1.0 + norm_b[inst]);
        const double rel_d = nrd / (1.0 + norm_c[inst]);
        const double rel_g = fabs(pobj - dobj) / (1.0 + fabs(pobj) + fabs(dobj));
        double k = rel_p > rel_d ? rel_p : rel_d;
        k = k > rel_g ? k : rel_g;
        if (isnan(rel_p) || isnan(rel_d) || isnan(rel_g)) k = NAN;
        double* o = out8 + 8 * inst;
        o[0] = k; o[1] = rel_p; o[2] = rel_d; o[3] = rel_g;
        o[4] = pobj; o[5] = dobj; o[6] = nrp; o[7] = nrd;
    }
}

// ---------------------------------------------------------------------------------
// Control: one workgroup per instance, run after the check passes of every block.
// Restart rule (reflected restarted Halpern PDHG, Lu & Yang): with r the weighted
// fixed-point residual sqrt(omega |dx|^2 + |dy|^2 / omega), restart when
// r <= 0.2 r0, or r <= 0.8 r0 and r > r_prev, or the epoch exceeds a fraction (default
// 0.5) of all iterations. With kkt_drift > 0 (by default FP32 values, and 16-bit values
// refined at every check) it also restarts when the FP64 rel-KKT of a check exceeds
// kkt_drift times the least rel-KKT of the epoch: the working-precision iteration then
// drifts towards the fixed point of its rounded data while the true error grows, and the
// restart (a new epoch from the current point) removes that error. At a restart the primal weight follows a proportional-integral
// controller on the error e = log(omega |Dx| / |Dy|) (|Dx|, |Dy|: the distances the epoch moved):
// log omega <- log omega - (theta e + ki sum(e)); with ki = 0 (the default, theta = 0.5) it is
// PDLP's smoothing log omega <- theta log(|Dy|/|Dx|) + (1 - theta) log omega.
// With refine_every_check the checks that do not restart also move the FP64 centre to the
// check point (iterative refinement every K iterations) and keep the Halpern epoch: the
// instance is then anchored, and its anchor is an explicit vector in the new coordinates.
struct ControlArgs {
    const double* px; int Rx;      // primal-step partials (cols), fixed tile height Rx
    const double* py; int Ry;      // dual-step partials (rows), fixed tile height Ry
    const int* xt_first;           // or per-instance tile ranges of variable-height tiles
    const int* xt_last;
    const int* yt_first;
    const int* yt_last;
    const double* kkt;             // out8 per instance from kkt_finish
    const int* row_off;
    const int* col_off;
    int K;
    double restart_sufficient, restart_necessary, restart_artificial, pw_theta;
    double kkt_drift;              // drift-restart factor, 0: off
    double pw_ki;                  // integral gain of the primal-weight controller
    int refine_every_check;        // 1: re-centre at every check, 0: at restarts only
    double* hist;                  // per instance: [checks] rel_kkt history (may be null)
    int hist_cap;
};

__global__ __launch_bounds__(kBlock) void control_kernel(InstState* __restrict__ st, ControlArgs a) {
    __shared__ double lds[kBlock];
    const int inst = blockIdx.x;
    if (st[inst].active == 0) return;
    const int64_t r_a = a.row_off[inst], r_b = a.row_off[inst + 1];
    const int64_t c_a = a.col_off[inst], c_b = a.col_off[inst + 1];
    double xq[3], yq[3];
    for (int q = 0; q < 3; ++q) {
        xq[q] = a.xt_first ? reduce_tiles(a.px, 3, q, inst, a.xt_first[inst], a.xt_last[inst], lds)
                           : reduce_instance(a.px, 3, q, inst, c_a, c_b, a.Rx, lds);
        yq[q] = a.yt_first ? reduce_tiles(a.py, 3, q, inst, a.yt_first[inst], a.yt_last[inst], lds)
                           : reduce_instance(a.py, 3, q, inst, r_a, r_b, a.Ry, lds);
    }
    const double xq0 = xq[0], xq1 = xq[1], xq2 = xq[2], yq0 = yq[0], yq1 = yq[1], yq2 = yq[2];
    if (threadIdx.x != 0) return;
    InstState s = st[inst];
    const double* k8 = a.kkt + 8 * inst;
    s.iters += a.K;
    s.k += a.K;
    s.rel_kkt = k8[0]; s.rel_p = k8[1]; s.rel_d = k8[2]; s.rel_gap = k8[3];
    s.pobj = k8[4]; s.dobj = k8[5];
    if (a.hist && s.checks < a.hist_cap) a.hist[static_cast<int64_t>(inst) * a.hist_cap + s.checks] = k8[0];
    s.checks += 1;
    s.restart_now = 0;
    const double r = sqrt(s.omega * xq1 + yq1 / s.omega);
    const double rf = sqrt(s.omega * xq0 + yq0 / s.omega);
    if (isnan(k8[0]) || isnan(r) || isnan(rf) || isinf(r)) {
        s.nan_flag = 1;
        s.active = 0;
    } else if (k8[0] <= s.tol) {
        s.converged = 1;
        s.active = 0;
    } else if (s.iters >= s.max_iters) {
        s.active = 0;
    } else {
        if (s.restarted_prev) {
            s.r0 = rf;
            s.r_prev = rf;
            s.kkt_min = k8[0];
        }
        if (s.kkt_min <= 0.0 || k8[0] < s.kkt_min) s.kkt_min = k8[0];
        const bool restart = (r <= a.restart_sufficient * s.r0) ||
                             (r <= a.restart_necessary * s.r0 && r > s.r_prev) ||
                             (static_cast<double>(s.k) >= a.restart_artificial * static_cast<double>(s.iters)) ||
                             (a.kkt_drift > 0.0 && k8[0] > a.kkt_drift * s.kkt_min);
        s.r_prev = r;
        s.restarted_prev = 0;
        s.anchored = (!restart && a.refine_every_check) ? 1 : 0;
        if (restart) {
            s.restart_now = 1;
            s.restarted_prev = 1;
            s.restarts += 1;
            s.k = 0;
            const double dx = sqrt(xq2), dy = sqrt(yq2);
            if (dx > 1e-10 && dy > 1e-10) {
                if (a.pw_ki == 0.0) {
                    s.omega = exp(a.pw_theta * log(dy / dx) + (1.0 - a.pw_theta) * log(s.omega));
                } else {
                    const double e = log(s.omega) - log(dy / dx);
                    s.pw_sum += e;
                    s.omega = exp(log(s.omega) - (a.pw_theta * e + a.pw_ki * s.pw_sum));
                }
                s.tau = s.eta / s.omega;
                s.sigma = s.eta * s.omega;
            }
        }
    }
    st[inst] = s;
}

// Restart (or initial centring, force = 1): move the centre to the last T(z) and
// rebuild the shifted data in the working precision. An anchored instance that did not
// restart moves its centre to the same point and keeps its epoch: the iterate and the anchor
// shift by d = T(z) - old centre (with rho = 1 the iterate is carried by xbar, x = lam xbar +
// pull an, so xbar shifts by d as well).
template <class T, bool MULTI>
__global__ void restart_cols(int n, const InstState* __restrict__ st, const int* __restrict__ col_inst,
                             int force, const double* __restrict__ xfull,
                             const double* __restrict__ aty_full, const double* __restrict__ c_s,
                             const double* __restrict__ lc_s, const double* __restrict__ uc_s,
                             double* __restrict__ xhat, T* __restrict__ x, T* __restrict__ xbar, T* __restrict__ ax,
                             T* __restrict__ dhat, T* __restrict__ lcs, T* __restrict__ ucs, double rho) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const int q = MULTI ? col_inst[j] : 0;
    const bool restart = force || st[q].restart_now != 0;
    if (!restart && !(st[q].anchored && st[q].active)) return;
    const double xf = xfull[j];
    if (restart) {
        x[j] = static_cast<T>(0);
        xbar[j] = static_cast<T>(0);
        if (ax) ax[j] = static_cast<T>(0);
    } else {
        const double d = xf - xhat[j];
        if (rho == 1.0) xbar[j] = static_cast<T>(static_cast<double>(xbar[j]) - d);
        else x[j] = static_cast<T>(static_cast<double>(x[j]) - d);
        ax[j] = static_cast<T>(static_cast<double>(ax[j]) - d);
    }
    xhat[j] = xf;
    dhat[j] = static_cast<T>(c_s[j] - aty_full[j]);
    lcs[j] = static_cast<T>(lc_s[j] - xf);
    ucs[j] = static_cast<T>(uc_s[j] - xf);
}

template <class T, bool MULTI>
__global__ void restart_rows(int m, const InstState* __restrict__ st, const int* __restrict__ row_inst,
                             int force, const double* __restrict__ yfull,
                             const double* __restrict__ ax_full, const double* __restrict__ lr_s,
                             const double* __restrict__ ur_s, double* __restrict__ yhat,
                             T* __restrict__ yhat_t, T* __restrict__ y, T* __restrict__ ay, T* __restrict__ lrs,
                             T* __restrict__ urs) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const int q = MULTI ? row_inst[i] : 0;
    const bool restart = force || st[q].restart_now != 0;
    if (!restart && !(st[q].anchored && st[q].active)) return;
    const double yf = yfull[i];
    if (restart) {
        y[i] = static_cast<T>(0);
        if (ay) ay[i] = static_cast<T>(0);
    } else {
        const double d = yf - yhat[i];
        y[i] = static_cast<T>(static_cast<double>(y[i]) - d);
        ay[i] = static_cast<T>(static_cast<double>(ay[i]) - d);
    }
    yhat[i] = yf;
    if (static_cast<void*>(yhat_t) != static_cast<void*>(yhat)) yhat_t[i] = static_cast<T>(yf);
    lrs[i] = static_cast<T>(lr_s[i] - ax_full[i]);
    urs[i] = static_cast<T>(ur_s[i] - ax_full[i]);
}

}  // namespace vj
