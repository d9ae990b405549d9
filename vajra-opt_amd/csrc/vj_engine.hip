// Restarted reflected-Halpern PDHG engine for one LP or a block-diagonal batch of LPs.
//
// Per iteration: primal_step (A^T y SpMV + projection + extrapolation + Halpern step)
// and dual_step (A x_bar SpMV + row prox + Halpern step). Every K iterations (one
// "block") the FP64 check runs: two KKT passes on the FP64 scaled matrix, the per-instance
// rel-KKT reduction, the control kernel (restart / primal weight / convergence, all on
// the device) and the conditional restart kernels. A block has fixed kernel arguments,
// so it is captured once as a hipGraph and replayed; the host only reads the per-instance
// state (a few hundred bytes) once per block, one block behind, into pinned memory.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <initializer_list>
#include <memory>
#include <vector>

#include "vj_context.hpp"
#include "vj_kernels.hpp"
#include "vj_layout.hpp"

namespace vj {

namespace {

using clk = std::chrono::steady_clock;
double secs(clk::time_point a, clk::time_point b) { return std::chrono::duration<double>(b - a).count(); }

// A HIP event owned by a scope, destroyed on every return path (error returns included).
struct ScopedEvent {
    hipEvent_t e = nullptr;
    ScopedEvent() = default;
    ScopedEvent(const ScopedEvent&) = delete;
    ScopedEvent& operator=(const ScopedEvent&) = delete;
    ~ScopedEvent() {
        if (e) (void)hipEventDestroy(e);
    }
};

#define VJ_DISPATCH_G(G_, ...)                         \
    switch (G_) {                                      \
        case 1: { constexpr int GG = 1; __VA_ARGS__; } break;   \
        case 2: { constexpr int GG = 2; __VA_ARGS__; } break;   \
        case 4: { constexpr int GG = 4; __VA_ARGS__; } break;   \
        case 8: { constexpr int GG = 8; __VA_ARGS__; } break;   \
        case 16: { constexpr int GG = 16; __VA_ARGS__; } break; \
        case 32: { constexpr int GG = 32; __VA_ARGS__; } break; \
        default: { set_error("unsupported lanes per row " + std::to_string(G_)); return VJ_ERR_INVALID_ARGUMENT; } \
    }

inline int tiles_for(int rows, int G) { return grid_for(rows, kBlock / G); }

uint16_t to_bf16_bits(double v) {
    float f = static_cast<float>(v);
    uint32_t u;
    std::memcpy(&u, &f, 4);
    if ((u & 0x7f800000u) == 0x7f800000u) return static_cast<uint16_t>(u >> 16);  // inf/nan
    const uint32_t lsb = (u >> 16) & 1u;
    u += 0x7fffu + lsb;  // round to nearest even
    return static_cast<uint16_t>(u >> 16);
}

__global__ void convert_f32(int64_t len, const double* __restrict__ a, float* __restrict__ b) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < len) b[i] = static_cast<float>(a[i]);
}
__global__ void convert_f16(int64_t len, const double* __restrict__ a, f16_t* __restrict__ b) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < len) b[i].bits = f16_bits_from(a[i]);
}
// part[b] = the largest |a| over the entries workgroup b visits (grid-stride, fixed order).
__global__ __launch_bounds__(kBlock) void max_abs_partial(int64_t len, const double* __restrict__ a,
                                                          double* __restrict__ part) {
    __shared__ double lds[kBlock];
    double m = 0.0;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x; i < len;
         i += static_cast<int64_t>(gridDim.x) * kBlock)
        m = fmax(m, fabs(a[i]));
    lds[threadIdx.x] = m;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if (static_cast<int>(threadIdx.x) < w) lds[threadIdx.x] = fmax(lds[threadIdx.x], lds[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = lds[0];
}
__global__ void convert_bf16(int64_t len, const double* __restrict__ a, bf16_t* __restrict__ b) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= len) return;
    float f = static_cast<float>(a[i]);
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7f800000u) != 0x7f800000u) u += 0x7fffu + ((u >> 16) & 1u);
    b[i].bits = static_cast<uint16_t>(u >> 16);
}

// c_s = C c ; lc_s = lc / C ; uc_s = uc / C (infinite bounds stay infinite)
__global__ void scale_cols_data(int n, const double* __restrict__ C, const double* __restrict__ c,
                                const double* __restrict__ lc, const double* __restrict__ uc,
                                double* __restrict__ c_s, double* __restrict__ lc_s,
                                double* __restrict__ uc_s) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const double s = C[j];
    c_s[j] = c[j] * s;
    lc_s[j] = lc[j] / s;
    uc_s[j] = uc[j] / s;
}
__global__ void scale_rows_data(int m, const double* __restrict__ R, const double* __restrict__ lr,
                                const double* __restrict__ ur, double* __restrict__ lr_s,
                                double* __restrict__ ur_s) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    lr_s[i] = lr[i] * R[i];
    ur_s[i] = ur[i] * R[i];
}
// per-instance objective/bound rescaling and the unscale factors
__global__ void rescale_cols(int n, const int* __restrict__ col_inst, const double* __restrict__ cs,
                             const double* __restrict__ bs, const double* __restrict__ C,
                             double* __restrict__ c_s, double* __restrict__ lc_s,
                             double* __restrict__ uc_s, double* __restrict__ ic,
                             double* __restrict__ px) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const int q = col_inst ? col_inst[j] : 0;
    c_s[j] *= cs[q];
    lc_s[j] *= bs[q];
    uc_s[j] *= bs[q];
    ic[j] = 1.0 / (cs[q] * C[j]);
    px[j] = C[j] / bs[q];
}
__global__ void rescale_rows(int m, const int* __restrict__ row_inst, const double* __restrict__ cs,
                             const double* __restrict__ bs, const double* __restrict__ R,
                             double* __restrict__ lr_s, double* __restrict__ ur_s,
                             double* __restrict__ ir, double* __restrict__ py) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const int q = row_inst ? row_inst[i] : 0;
    lr_s[i] *= bs[q];
    ur_s[i] *= bs[q];
    ir[i] = 1.0 / (bs[q] * R[i]);
    py[i] = R[i] / cs[q];
}
// x0 = clamp(0, lc, uc)
__global__ void initial_point(int n, const double* __restrict__ lc, const double* __restrict__ uc,
                              double* __restrict__ x) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    double v = 0.0;
    if (v < lc[j]) v = lc[j];
    if (v > uc[j]) v = uc[j];
    x[j] = v;
}
// v = u / sqrt(nrm2[inst])
__global__ void normalize_seg(int n, const int* __restrict__ inst_of, const double* __restrict__ nrm2,
                              const double* __restrict__ u, double* __restrict__ v) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const int q = inst_of ? inst_of[j] : 0;
    const double s = nrm2[q];
    v[j] = s > 0.0 ? u[j] / sqrt(s) : 0.0;
}
__global__ void init_unit_seg(int n, const int* __restrict__ inst_of, const int* __restrict__ off,
                              double* __restrict__ v) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const int q = inst_of ? inst_of[j] : 0;
    v[j] = 1.0 / sqrt(static_cast<double>(off[q + 1] - off[q]));
}
// original-space solution: x = xfull * px ; y_canonical = yfull * py
__global__ void unscale_out(int len, const double* __restrict__ a, const double* __restrict__ f,
                            double* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < len) out[i] = a[i] * f[i];
}

}  // namespace

// The SpMV layout of one fused step, as the autotuner picks it.
struct StepCfg {
    int S, G;
    bool stream;
    bool rec;   // FP32 (index, value) records in the tile-streamed kernels
    bool ell;   // sliced-ELL kernel (S = 1)
    bool stage = false;   // LDS-staged kernel (S = 1)
    bool w16 = false;     // ... reading 16-bit window offsets
    int shape = 0;        // ... in workgroups of kStageShapeThreads[shape] threads
};

// An autotuner decision, kept for the rest of the upload: the same precision and kernel
// parameters on the same problem reuse it, so repeated default solves run the same kernels
// (and, the kernels being deterministic, return the same bits).
struct TuneDecision {
    int prec, lanes_A, lanes_At, col_splits_A, col_splits_At, check_frequency, stage_threads;
    StepCfg a, at;
    int GA, GAt, K_tuned;
};

// Row tiles of the LDS-staged kernels for one step and workgroup shape: tiles that never span
// two instances, at most rpl * threads rows and cap entries (cap: the entry capacity they were
// built for, 0 while none are; it differs between 8-byte and 4-byte iterates), each tile's
// column window (win), per instance the first / last tile (tf / tl), and the sum of the
// windows' widths (gwin). With every window at most 65,536 columns wide (w16ok), 16-bit window
// offsets of the entries (off16) replace the 4-byte column indices, and BF16/FP16 values come
// as 4-byte (offset, value) records (w16rec). inst: each tile's instance (hInst: on the host);
// desc: per tile rows, entries, window and instance (8 ints). list: the tiles a staged launch
// runs, nlist of them (its grid): every tile at the start of a solve, then the tiles of the
// instances still active, rebuilt by the host (with the stream drained) when half of the
// listed tiles belong to finished instances.
struct StageLayout {
    DevBuf tile, win, tf, tl, off16, inst, list, desc, w16rec;
    std::vector<int> hInst;
    int threads = 0, rpl = 1, nt = 0, nlist = 0, cap = 0;
    int64_t gwin = 0;
    bool w16ok = false;
};
// Threads per workgroup of the staged steps, per shape index.
constexpr int kStageShapeThreads[2] = {1024, 256};

struct FastEngine {
    int prec = VJ_PREC_FP64;
    int count = 0, m = 0, n = 0;
    int64_t nnz = 0;
    bool multi = false;
    int GA = 8, GAt = 4, K = 64;
    int K_tuned = 0;   // iterations per check picked by the autotuner (0: not timed)
    std::vector<TuneDecision> tuned;   // decisions of earlier solves of this upload
    double rho = 1.0;
    DevBuf av_s, tv_s, av_lo, tv_lo;
    DevBuf Rsc, Csc, rstep, cstep, rnorm, cnorm;
    DevBuf ir, ic, pxf, pyf;
    DevBuf c_s, lc_s, uc_s, lr_s, ur_s;
    DevBuf x, xbar, y, dhat, lcs, ucs, lrs, urs, yhat_t;
    DevBuf ax, ay;   // Halpern anchors of anchored instances (the centre moved inside the epoch)
    DevBuf xhat, yhat, xfull, yfull, ax_full, aty_full;
    DevBuf part_x, part_y, part_kr, part_kc, part_s;
    DevBuf kkt8, st, offset_d, norm_b_d, norm_c_d, objfac_d, inst_tmp_a, inst_tmp_b, hist;
    DevBuf vtmp_n, wtmp_m, utmp_n;
    // XCD-split layouts (single LP): A by column blocks (gathers x_bar), A^T by row blocks
    // of A (gathers y); S = 1 means the fused single-pass kernels are used
    int SA = 1, SAt = 1, GsA = 2, GsAt = 2;
    DevBuf bpA, idxA, valA, psumA, bpAt, idxAt, valAt, psumAt, permA, permAt;
    // row tiles of the split layouts for split_stream_partials (Gs = 0 selects that kernel),
    // each tile's first entry in every block, tbase[t * S + k] (t = 0..tiles), and each row's
    // first entry in every block relative to its tile's (toff*, 16-bit)
    DevBuf tileSA, tileSAt, tbaseSA, tbaseSAt, toffSA, toffSAt;
    int ntSA = 0, ntSAt = 0, tcapSA = 0, tcapSAt = 0;   // tile count; LDS entries the tiles fit
    // FP32: interleaved (index, value) records of A, A^T and of the split layouts, read by
    // the tile-streamed kernels when the autotuner picks them (recA / recAt)
    DevBuf recA_buf, recAt_buf, recSA_buf, recSAt_buf;
    bool recA = false, recAt = false;
    // sliced-ELL copies (single LP): per-slice entry offsets and widths, entries in slice
    // order, their CSR positions (-1 = padding), values and FP32 records; ellA / ellAt select
    // the ELL step kernels; ellPad* = ELL entries / nonzeros (0 while not measured)
    DevBuf ellOffA, ellWA, ellIdxA, ellPermA, ellValA, ellRecA;
    DevBuf ellOffAt, ellWAt, ellIdxAt, ellPermAt, ellValAt, ellRecAt;
    int64_t ellLenA = 0, ellLenAt = 0;
    bool ellBuiltA = false, ellBuiltAt = false, ellA = false, ellAt = false;
    double ellPadA = 0.0, ellPadAt = 0.0;
    int builtSA = 0, builtSAt = 0;   // split layout currently held in bp/idx/perm
    // tile-streamed kernels: host-built row tiles of A (y side) and A^T (x side) and, per
    // instance, the first/last tile touching it
    bool streamA = true, streamAt = true, tiles_built = false;
    int ntA = 0, ntAt = 0;
    DevBuf tileA, tileAt, tfA, tlA, tfAt, tlAt;
    // batches: the tile-streamed launches' lists of tiles holding a row of an active instance
    // (listTA / listTAt, nlistT* of them), kept like the staged lists; hFirstT* / hLastT*: the
    // first and last instance of each tile, on the host
    DevBuf listTA, listTAt;
    std::vector<int> hFirstTA, hLastTA, hFirstTAt, hLastTAt;
    int nlistTA = 0, nlistTAt = 0;
    // LDS-staged kernels (batches), one StageLayout per step and workgroup shape (stA / stAt
    // [0]: 1,024 threads, [1]: 256; shA / shAt: the shape the steps run, LA() / LAt() its
    // layout); stageA / stageAt select the staged steps, w16A / w16At their 16-bit window
    // offsets. pruned: some list leaves tiles out.
    StageLayout stA[2], stAt[2];
    int shA = 0, shAt = 0;
    StageLayout& LA() { return stA[shA]; }
    StageLayout& LAt() { return stAt[shAt]; }
    bool pruned = false;
    bool stageA = false, stageAt = false, w16A = false, w16At = false;
    int hist_cap = 0;
    hipGraphExec_t gexec = nullptr;
    InstState* h_state[2] = {nullptr, nullptr};   // pinned, h_state_cap entries each
    int h_state_cap = 0;
    std::vector<double> spectral;
    std::vector<double> h_hist;
    ~FastEngine() {
        if (gexec) (void)hipGraphExecDestroy(gexec);
        for (auto* p : h_state)
            if (p) (void)hipHostFree(p);
    }
    size_t tsize() const { return prec == VJ_PREC_FP64 ? sizeof(double) : sizeof(float); }
};

void destroy_fast(FastEngine* e) { delete e; }

namespace {

// enough slots for the largest lane count (32 lanes: 8 rows per tile)
int partial_slots(int rows, int count) { return tiles_for(rows, 32) + count + 1; }

// ---- launch helpers ---------------------------------------------------------------
int require(std::initializer_list<const void*> ptrs, const char* where);

// Call go(entries) with the entry accessor the chosen variant reads: FP32 records when
// `rec` is set (FP32 values only), else the separate index / value arrays.
template <class V, class F>
int with_entries(bool rec, const DevBuf& recs, const int* idx, const V* val, const char* where, F go) {
    if constexpr (std::is_same<V, float>::value) {
        if (rec) {
            VJ_TRY(require({recs.p}, where));
            go(RecEntries{recs.as<vj_i32x2>()});
            return VJ_OK;
        }
    }
    go(SepEntries<V>{idx, val});
    return VJ_OK;
}

template <class V, class T>
int launch_primal(vj_ctx* ctx, FastEngine& E, const V* tv, T* x, T* xbar, const T* y, int it, int flags) {
    const T* ax = E.ax.as<T>();
    Problem& P = ctx->prob;
    const int tiles = tiles_for(E.n, E.GAt);
    VJ_TRY(require({P.tp.p, P.ti.p, tv, y, E.dhat.p, E.lcs.p, E.ucs.p, x, xbar, E.xhat.p, E.xfull.p, E.st.p,
                    E.part_x.p, E.multi ? P.col_inst.p : E.st.p}, "primal_step"));
    if (E.SAt > 1) {
        VJ_TRY(require({E.bpAt.p, E.idxAt.p, E.valAt.p, E.psumAt.p}, "split primal_step"));
        if (E.GsAt == 0) {
            VJ_TRY(require({E.tileSAt.p, E.tbaseSAt.p, E.toffSAt.p}, "split_stream_partials"));
            VJ_TRY(with_entries<V>(E.recAt, E.recSAt_buf, E.idxAt.as<int>(), E.valAt.as<V>(), "split records",
                                   [&](auto ent) {
                hipLaunchKernelGGL((split_stream_partials<decltype(ent), T>), dim3(E.SAt * E.ntSAt), dim3(kBlock), 0,
                                   ctx->stream, E.n, E.SAt, E.tileSAt.as<int>(), E.tbaseSAt.as<int>(),
                                   E.toffSAt.as<uint16_t>(), ent, y, E.psumAt.as<T>(), E.st.as<InstState>());
            }));
        } else {
            VJ_DISPATCH_G(E.GsAt, hipLaunchKernelGGL((split_partials<V, T, GG>), dim3(E.SAt * tiles_for(E.n, GG)),
                dim3(kBlock), 0, ctx->stream, E.n, E.SAt, E.bpAt.as<int>(), E.idxAt.as<int>(), E.valAt.as<V>(), y,
                E.psumAt.as<T>(), E.st.as<InstState>()));
        }
        hipLaunchKernelGGL((primal_finish<T>), dim3(grid_for(E.n, kBlock * finish_rows<T>())), dim3(kBlock), 0, ctx->stream, E.n, E.SAt,
                           E.psumAt.as<T>(), E.dhat.as<T>(), E.lcs.as<T>(), E.ucs.as<T>(), x, xbar, ax, E.xhat.as<double>(),
                           E.xfull.as<double>(), E.st.as<InstState>(), E.part_x.as<double>(), it, flags, E.rho);
        VJ_LAUNCHED();
        return VJ_OK;
    }
    if (E.ellAt) {
        VJ_TRY(require({E.ellOffAt.p, E.ellWAt.p, E.ellIdxAt.p, E.ellValAt.p}, "ell primal_step"));
        VJ_TRY(with_entries<V>(E.recAt, E.ellRecAt, E.ellIdxAt.as<int>(), E.ellValAt.as<V>(), "ell records",
                               [&](auto ent) {
            hipLaunchKernelGGL((ell_primal_step<decltype(ent), T>), dim3(grid_for(E.n, kBlock)), dim3(kBlock), 0,
                               ctx->stream, E.n, E.ellOffAt.as<int>(), E.ellWAt.as<int>(), ent, y, E.dhat.as<T>(),
                               E.lcs.as<T>(), E.ucs.as<T>(), x, xbar, ax, E.xhat.as<double>(), E.xfull.as<double>(),
                               E.st.as<InstState>(), E.part_x.as<double>(), it, flags, E.rho);
        }));
        VJ_LAUNCHED();
        return VJ_OK;
    }
    if (E.stageAt) {
        VJ_TRY(require({E.LAt().desc.p, E.LAt().list.p}, "staged primal_step"));
        if (E.LAt().nlist == 0) return VJ_OK;   // every instance has finished
        auto go = [&](auto ent) {
            using Ent = decltype(ent);
            auto go2 = [&](auto multi, auto rpl) {
                auto go3 = [&](auto nt) {
                hipLaunchKernelGGL((staged_primal_step<Ent, T, decltype(multi)::value, decltype(rpl)::value,
                                                       decltype(nt)::value>),
                                   dim3(E.LAt().nlist), dim3(decltype(nt)::value), 0, ctx->stream, E.LAt().list.as<int>(),
                                   E.LAt().desc.as<int4>(), P.tp.as<int>(), ent, y,
                                   E.dhat.as<T>(), E.lcs.as<T>(), E.ucs.as<T>(), x, xbar, ax, E.xhat.as<double>(),
                                   E.xfull.as<double>(), E.st.as<InstState>(), E.part_x.as<double>(), it, flags, E.rho);
                };
                if (E.LAt().threads == 256) go3(std::integral_constant<int, 256>{});
                else go3(std::integral_constant<int, 1024>{});
            };
            using One = std::integral_constant<int, 1>;
            using Two = std::integral_constant<int, 2>;
            if (E.multi) {
                if (E.LAt().rpl == 2) go2(std::true_type{}, Two{}); else go2(std::true_type{}, One{});
            } else {
                if (E.LAt().rpl == 2) go2(std::false_type{}, Two{}); else go2(std::false_type{}, One{});
            }
        };
        if (E.w16At) {
            VJ_TRY(require({E.LAt().off16.p}, "staged primal_step (16-bit offsets)"));
            if constexpr (std::is_same<V, bf16_t>::value || std::is_same<V, f16_t>::value) {
                VJ_TRY(require({E.LAt().w16rec.p}, "staged primal_step (16-bit value records)"));
                go(Win16Records<V>{E.LAt().w16rec.as<uint32_t>()});
            } else {
                go(Win16Entries<V>{E.LAt().off16.as<uint16_t>(), tv});
            }
        } else {
            VJ_TRY(with_entries<V>(E.recAt, E.recAt_buf, P.ti.as<int>(), tv, "staged records", go));
        }
        VJ_LAUNCHED();
        return VJ_OK;
    }
    if (E.streamAt) {
        VJ_TRY(require({E.tileAt.p}, "stream primal_step"));
        VJ_TRY(with_entries<V>(E.recAt, E.recAt_buf, P.ti.as<int>(), tv, "stream records", [&](auto ent) {
            using Ent = decltype(ent);
            if (E.multi) {
                if (E.nlistTAt > 0)
                hipLaunchKernelGGL((stream_primal_step<Ent, T, true>), dim3(E.nlistTAt), dim3(kBlock), 0, ctx->stream, E.n,
                                   E.listTAt.as<int>(), E.tileAt.as<int>(), P.tp.as<int>(), ent, y, E.dhat.as<T>(), E.lcs.as<T>(),
                                   E.ucs.as<T>(), x, xbar, ax, E.xhat.as<double>(), E.xfull.as<double>(),
                                   E.st.as<InstState>(), P.col_inst.as<int>(), E.part_x.as<double>(), it, flags, E.rho);
            } else
                hipLaunchKernelGGL((stream_primal_step<Ent, T, false>), dim3(E.ntAt), dim3(kBlock), 0, ctx->stream, E.n,
                                   nullptr, E.tileAt.as<int>(), P.tp.as<int>(), ent, y, E.dhat.as<T>(), E.lcs.as<T>(),
                                   E.ucs.as<T>(), x, xbar, ax, E.xhat.as<double>(), E.xfull.as<double>(),
                                   E.st.as<InstState>(), P.col_inst.as<int>(), E.part_x.as<double>(), it, flags, E.rho);
        }));
        VJ_LAUNCHED();
        return VJ_OK;
    }
    if (E.multi) {
        VJ_DISPATCH_G(E.GAt, hipLaunchKernelGGL((primal_step<V, T, GG, true>), dim3(tiles), dim3(kBlock), 0, ctx->stream,
            E.n, P.tp.as<int>(), P.ti.as<int>(), tv, y, E.dhat.as<T>(), E.lcs.as<T>(), E.ucs.as<T>(), x, xbar, ax,
            E.xhat.as<double>(), E.xfull.as<double>(), E.st.as<InstState>(), P.col_inst.as<int>(),
            E.part_x.as<double>(), it, flags, E.rho));
    } else {
        VJ_DISPATCH_G(E.GAt, hipLaunchKernelGGL((primal_step<V, T, GG, false>), dim3(tiles), dim3(kBlock), 0, ctx->stream,
            E.n, P.tp.as<int>(), P.ti.as<int>(), tv, y, E.dhat.as<T>(), E.lcs.as<T>(), E.ucs.as<T>(), x, xbar, ax,
            E.xhat.as<double>(), E.xfull.as<double>(), E.st.as<InstState>(), P.col_inst.as<int>(),
            E.part_x.as<double>(), it, flags, E.rho));
    }
    VJ_LAUNCHED();
    return VJ_OK;
}

// FP64 iterates read the FP64 centre directly; FP32/BF16 read its working-precision copy.
template <class T>
T* yhat_working(FastEngine& E) {
    return E.prec == VJ_PREC_FP64 ? reinterpret_cast<T*>(E.yhat.p) : E.yhat_t.as<T>();
}

// Every device pointer a kernel dereferences must be live: refuse to launch otherwise.
int require(std::initializer_list<const void*> ptrs, const char* where) {
    for (const void* q : ptrs)
        if (!q) {
            set_error(std::string("internal: null device buffer passed to ") + where);
            return VJ_ERR_LAUNCH;
        }
    return VJ_OK;
}

template <class V, class T>
int launch_dual(vj_ctx* ctx, FastEngine& E, const V* av, const T* xbar, T* y, int it, int flags) {
    const T* ay = E.ay.as<T>();
    Problem& P = ctx->prob;
    const int tiles = tiles_for(E.m, E.GA);
    const T* yht = yhat_working<T>(E);
    // all rows equalities: the shifted bounds are equal, and passing one vector for both
    // selects the kernels' equality-row update (no upper bound, no centre dual read)
    const T* lrs = E.lrs.as<T>();
    const T* urs = P.all_eq ? lrs : E.urs.as<T>();
    VJ_TRY(require({P.rp.p, P.ci.p, av, xbar, yht, E.lrs.p, E.urs.p, y, E.yhat.p, E.yfull.p, E.st.p, E.part_y.p,
                    E.multi ? P.row_inst.p : E.st.p}, "dual_step"));
    if (E.SA > 1) {
        VJ_TRY(require({E.bpA.p, E.idxA.p, E.valA.p, E.psumA.p}, "split dual_step"));
        if (E.GsA == 0) {
            VJ_TRY(require({E.tileSA.p, E.tbaseSA.p, E.toffSA.p}, "split_stream_partials"));
            VJ_TRY(with_entries<V>(E.recA, E.recSA_buf, E.idxA.as<int>(), E.valA.as<V>(), "split records",
                                   [&](auto ent) {
                hipLaunchKernelGGL((split_stream_partials<decltype(ent), T>), dim3(E.SA * E.ntSA), dim3(kBlock), 0,
                                   ctx->stream, E.m, E.SA, E.tileSA.as<int>(), E.tbaseSA.as<int>(),
                                   E.toffSA.as<uint16_t>(), ent, xbar, E.psumA.as<T>(), E.st.as<InstState>());
            }));
        } else {
            VJ_DISPATCH_G(E.GsA, hipLaunchKernelGGL((split_partials<V, T, GG>), dim3(E.SA * tiles_for(E.m, GG)),
                dim3(kBlock), 0, ctx->stream, E.m, E.SA, E.bpA.as<int>(), E.idxA.as<int>(), E.valA.as<V>(), xbar,
                E.psumA.as<T>(), E.st.as<InstState>()));
        }
        hipLaunchKernelGGL((dual_finish<T>), dim3(grid_for(E.m, kBlock * finish_rows<T>())), dim3(kBlock), 0, ctx->stream, E.m, E.SA,
                           E.psumA.as<T>(), yht, lrs, urs, y, ay, E.yhat.as<double>(),
                           E.yfull.as<double>(), E.st.as<InstState>(), E.part_y.as<double>(), it, flags, E.rho);
        VJ_LAUNCHED();
        return VJ_OK;
    }
    if (E.ellA) {
        VJ_TRY(require({E.ellOffA.p, E.ellWA.p, E.ellIdxA.p, E.ellValA.p}, "ell dual_step"));
        VJ_TRY(with_entries<V>(E.recA, E.ellRecA, E.ellIdxA.as<int>(), E.ellValA.as<V>(), "ell records",
                               [&](auto ent) {
            hipLaunchKernelGGL((ell_dual_step<decltype(ent), T>), dim3(grid_for(E.m, kBlock)), dim3(kBlock), 0,
                               ctx->stream, E.m, E.ellOffA.as<int>(), E.ellWA.as<int>(), ent, xbar, yht,
                               lrs, urs, y, ay, E.yhat.as<double>(), E.yfull.as<double>(),
                               E.st.as<InstState>(), E.part_y.as<double>(), it, flags, E.rho);
        }));
        VJ_LAUNCHED();
        return VJ_OK;
    }
    if (E.stageA) {
        VJ_TRY(require({E.LA().desc.p, E.LA().list.p}, "staged dual_step"));
        if (E.LA().nlist == 0) return VJ_OK;
        auto go = [&](auto ent) {
            using Ent = decltype(ent);
            auto go2 = [&](auto multi, auto rpl) {
                auto go3 = [&](auto nt) {
                hipLaunchKernelGGL((staged_dual_step<Ent, T, decltype(multi)::value, decltype(rpl)::value,
                                                     decltype(nt)::value>),
                                   dim3(E.LA().nlist), dim3(decltype(nt)::value), 0, ctx->stream, E.LA().list.as<int>(),
                                   E.LA().desc.as<int4>(), P.rp.as<int>(), ent, xbar,
                                   yht, lrs, urs, y, ay, E.yhat.as<double>(), E.yfull.as<double>(), E.st.as<InstState>(),
                                   E.part_y.as<double>(), it, flags, E.rho);
                };
                if (E.LA().threads == 256) go3(std::integral_constant<int, 256>{});
                else go3(std::integral_constant<int, 1024>{});
            };
            using One = std::integral_constant<int, 1>;
            using Two = std::integral_constant<int, 2>;
            if (E.multi) {
                if (E.LA().rpl == 2) go2(std::true_type{}, Two{}); else go2(std::true_type{}, One{});
            } else {
                if (E.LA().rpl == 2) go2(std::false_type{}, Two{}); else go2(std::false_type{}, One{});
            }
        };
        if (E.w16A) {
            VJ_TRY(require({E.LA().off16.p}, "staged dual_step (16-bit offsets)"));
            if constexpr (std::is_same<V, bf16_t>::value || std::is_same<V, f16_t>::value) {
                VJ_TRY(require({E.LA().w16rec.p}, "staged dual_step (16-bit value records)"));
                go(Win16Records<V>{E.LA().w16rec.as<uint32_t>()});
            } else {
                go(Win16Entries<V>{E.LA().off16.as<uint16_t>(), av});
            }
        } else {
            VJ_TRY(with_entries<V>(E.recA, E.recA_buf, P.ci.as<int>(), av, "staged records", go));
        }
        VJ_LAUNCHED();
        return VJ_OK;
    }
    if (E.streamA) {
        VJ_TRY(require({E.tileA.p}, "stream dual_step"));
        VJ_TRY(with_entries<V>(E.recA, E.recA_buf, P.ci.as<int>(), av, "stream records", [&](auto ent) {
            using Ent = decltype(ent);
            if (E.multi) {
                if (E.nlistTA > 0)
                hipLaunchKernelGGL((stream_dual_step<Ent, T, true>), dim3(E.nlistTA), dim3(kBlock), 0, ctx->stream, E.m,
                                   E.listTA.as<int>(), E.tileA.as<int>(), P.rp.as<int>(), ent, xbar, yht, lrs, urs, y, ay,
                                   E.yhat.as<double>(), E.yfull.as<double>(), E.st.as<InstState>(),
                                   P.row_inst.as<int>(), E.part_y.as<double>(), it, flags, E.rho);
            } else
                hipLaunchKernelGGL((stream_dual_step<Ent, T, false>), dim3(E.ntA), dim3(kBlock), 0, ctx->stream, E.m,
                                   nullptr, E.tileA.as<int>(), P.rp.as<int>(), ent, xbar, yht, lrs, urs, y, ay,
                                   E.yhat.as<double>(), E.yfull.as<double>(), E.st.as<InstState>(),
                                   P.row_inst.as<int>(), E.part_y.as<double>(), it, flags, E.rho);
        }));
        VJ_LAUNCHED();
        return VJ_OK;
    }
    if (E.multi) {
        VJ_DISPATCH_G(E.GA, hipLaunchKernelGGL((dual_step<V, T, GG, true>), dim3(tiles), dim3(kBlock), 0, ctx->stream,
            E.m, P.rp.as<int>(), P.ci.as<int>(), av, xbar, yht, lrs, urs, y, ay,
            E.yhat.as<double>(), E.yfull.as<double>(), E.st.as<InstState>(), P.row_inst.as<int>(),
            E.part_y.as<double>(), it, flags, E.rho));
    } else {
        VJ_DISPATCH_G(E.GA, hipLaunchKernelGGL((dual_step<V, T, GG, false>), dim3(tiles), dim3(kBlock), 0, ctx->stream,
            E.m, P.rp.as<int>(), P.ci.as<int>(), av, xbar, yht, lrs, urs, y, ay,
            E.yhat.as<double>(), E.yfull.as<double>(), E.st.as<InstState>(), P.row_inst.as<int>(),
            E.part_y.as<double>(), it, flags, E.rho));
    }
    VJ_LAUNCHED();
    return VJ_OK;
}

// KKT passes + finish. scaled=true: in-loop check on the scaled problem (writes ax_full /
// aty_full, unscale arrays applied); scaled=false: audit on original data.
int launch_kkt(vj_ctx* ctx, FastEngine* E, int GA, int GAt, bool multi, const double* x,
               const double* y, bool scaled, int only_active, double* part_r, double* part_c,
               double* out8) {
    Problem& P = ctx->prob;
    const int m = P.m, n = P.n;
    const double* av = scaled ? E->av_s.as<double>() : P.av.as<double>();
    const double* tv = scaled ? E->tv_s.as<double>() : P.tv.as<double>();
    const double* lr = scaled ? E->lr_s.as<double>() : P.lr.as<double>();
    const double* ur = scaled ? E->ur_s.as<double>() : P.ur.as<double>();
    const double* lc = scaled ? E->lc_s.as<double>() : P.lc.as<double>();
    const double* uc = scaled ? E->uc_s.as<double>() : P.uc.as<double>();
    const double* c = scaled ? E->c_s.as<double>() : P.c_can.as<double>();
    const double* ir = scaled ? E->ir.as<double>() : nullptr;
    const double* ic = scaled ? E->ic.as<double>() : nullptr;
    const double* py = scaled ? E->pyf.as<double>() : nullptr;
    double* axo = scaled ? E->ax_full.as<double>() : nullptr;
    double* atyo = scaled ? E->aty_full.as<double>() : nullptr;
    const InstState* st = E ? E->st.as<InstState>() : nullptr;
    const int tr = tiles_for(m, GA), tc = tiles_for(n, GAt);
    VJ_TRY(require({P.rp.p, P.ci.p, av, x, y, lr, ur, P.tp.p, P.ti.p, tv, c, lc, uc, part_r, part_c, out8,
                    P.d_row_off.p, P.d_col_off.p, E->norm_b_d.p, E->norm_c_d.p, E->offset_d.p,
                    multi ? P.row_inst.p : part_r, multi ? P.col_inst.p : part_r,
                    only_active ? static_cast<const void*>(st) : part_r},
                   "kkt passes"));
    if (scaled) VJ_TRY(require({ir, ic, py, axo, atyo, E->objfac_d.p}, "kkt passes (scaled)"));
    if (multi) {
        VJ_DISPATCH_G(GA, hipLaunchKernelGGL((kkt_rows<GG, true>), dim3(tr), dim3(kBlock), 0, ctx->stream, m,
            P.rp.as<int>(), P.ci.as<int>(), av, x, y, lr, ur, ir, py, axo, st, only_active, P.row_inst.as<int>(), part_r));
        VJ_DISPATCH_G(GAt, hipLaunchKernelGGL((kkt_cols<GG, true>), dim3(tc), dim3(kBlock), 0, ctx->stream, n,
            P.tp.as<int>(), P.ti.as<int>(), tv, y, x, c, lc, uc, ic, atyo, st, only_active, P.col_inst.as<int>(), part_c));
    } else {
        VJ_DISPATCH_G(GA, hipLaunchKernelGGL((kkt_rows<GG, false>), dim3(tr), dim3(kBlock), 0, ctx->stream, m,
            P.rp.as<int>(), P.ci.as<int>(), av, x, y, lr, ur, ir, py, axo, st, only_active, P.row_inst.as<int>(), part_r));
        VJ_DISPATCH_G(GAt, hipLaunchKernelGGL((kkt_cols<GG, false>), dim3(tc), dim3(kBlock), 0, ctx->stream, n,
            P.tp.as<int>(), P.ti.as<int>(), tv, y, x, c, lc, uc, ic, atyo, st, only_active, P.col_inst.as<int>(), part_c));
    }
    VJ_LAUNCHED();
    const double* objfac = scaled ? E->objfac_d.as<double>() : nullptr;
    hipLaunchKernelGGL(kkt_finish, dim3(P.count), dim3(kBlock), 0, ctx->stream, part_r, kBlock / GA, part_c,
                       kBlock / GAt, P.d_row_off.as<int>(), P.d_col_off.as<int>(), E ? E->norm_b_d.as<double>() : nullptr,
                       E ? E->norm_c_d.as<double>() : nullptr, E ? E->offset_d.as<double>() : nullptr, objfac, out8);
    VJ_LAUNCHED();
    return VJ_OK;
}

// Per-instance sum of squares of a vector (mode 0) or of the finite-bound vector (mode 1).
int seg_norm2(vj_ctx* ctx, FastEngine& E, int len, int mode, const double* v, const double* lo,
              const double* hi, bool rows, double* out, const double* skip = nullptr) {
    Problem& P = ctx->prob;
    const int tiles = grid_for(len, kBlock);
    const int* inst = rows ? P.row_inst.as<int>() : P.col_inst.as<int>();
    if (E.multi)
        hipLaunchKernelGGL((seg_sumsq<true>), dim3(tiles), dim3(kBlock), 0, ctx->stream, len, mode, v, lo, hi, skip,
                           inst, E.part_s.as<double>());
    else
        hipLaunchKernelGGL((seg_sumsq<false>), dim3(tiles), dim3(kBlock), 0, ctx->stream, len, mode, v, lo, hi, skip,
                           inst, E.part_s.as<double>());
    VJ_LAUNCHED();
    hipLaunchKernelGGL(seg_finish, dim3(P.count), dim3(kBlock), 0, ctx->stream, E.part_s.as<double>(), 1, 0,
                       rows ? P.d_row_off.as<int>() : P.d_col_off.as<int>(), kBlock, out);
    VJ_LAUNCHED();
    return VJ_OK;
}

int spmv_f64(vj_ctx* ctx, int rows, const int* rp, const int* ci, const double* av, const double* x, double* y, int G) {
    VJ_DISPATCH_G(G, hipLaunchKernelGGL((spmv_rows_f64<GG>), dim3(tiles_for(rows, GG)), dim3(kBlock), 0, ctx->stream,
                                        rows, rp, ci, av, x, y));
    VJ_LAUNCHED();
    return VJ_OK;
}

int choose_lanes(double avg_nnz) {
    if (avg_nnz <= 3.0) return 2;
    if (avg_nnz <= 8.0) return 4;
    if (avg_nnz <= 24.0) return 8;
    if (avg_nnz <= 64.0) return 16;
    return 32;
}

// Column blocks of at most 512K entries of the gathered vector (2 MB in FP32, 4 MB in FP64):
// on the 1M x 2M LP this block size was the fastest split for both precisions (S = 2, 4, 8
// measured; FP64 at S = 4 beat S = 8 by 6%, FP32 at S = 4 beat S = 2 and S = 8).
int choose_splits(int64_t vec_len) {
    const int64_t target = 512ll << 10;
    int S = 1;
    while (S < 8 && vec_len > target * S) S *= 2;
    return S;
}

// first[t] / last[t]: the first and last instance holding a row of tile t (rows off[q] ..
// off[q+1]-1 belong to instance q), from the device tile array.
int tile_instance_ranges(hipStream_t s, int ntiles, const DevBuf& tile_row, const std::vector<int64_t>& off,
                         std::vector<int>& first, std::vector<int>& last) {
    std::vector<int> tr(static_cast<size_t>(ntiles) + 1);
    VJ_HIP(hipMemcpyAsync(tr.data(), tile_row.p, tr.size() * 4, hipMemcpyDeviceToHost, s));
    VJ_HIP(hipStreamSynchronize(s));
    auto inst_of = [&](int64_t row) {   // the last q with off[q] <= row
        return static_cast<int>(std::upper_bound(off.begin(), off.end(), row) - off.begin()) - 1;
    };
    first.resize(ntiles);
    last.resize(ntiles);
    for (int t = 0; t < ntiles; ++t) {
        first[t] = inst_of(tr[t]);
        last[t] = inst_of(tr[t + 1] - 1);
    }
    return VJ_OK;
}

// Rows per lane of a staged tile: 2 where one row per thread of the average length fills at
// most three quarters of the tile's entry capacity (short rows: the columns of a random
// instance, 6 entries, or of a staircase instance, about 8), else 1. Tiles still stop at the
// capacity.
int stage_rows_per_lane(double avg_row, int cap, int threads) { return threads * avg_row <= 0.75 * cap ? 2 : 1; }

// The staged tiles of one step (rows rp / indices ci, `rows` rows) for one workgroup shape,
// built on the device, with every tile listed.
int build_stage_layout(hipStream_t s, FastEngine& E, StageLayout& L, int threads, int rows, const int* rp,
                       const int* ci, const int* inst_of, const int* off, const int* d_off) {
    const int cap = threads * kStageLaneBytes / static_cast<int>(E.prec == VJ_PREC_FP64 ? 8 : 4);
    int widest = 0;
    L.threads = threads;
    L.rpl = stage_rows_per_lane(static_cast<double>(E.nnz) / std::max(1, rows), cap, threads);
    VJ_TRY(device_row_tiles(s, rows, 1, rp, 0, cap, L.rpl * threads, L.tile, &L.nt, inst_of, off));
    VJ_TRY(device_tile_windows(s, L.nt, L.tile.as<int>(), rp, ci, L.win, &L.gwin, &widest));
    VJ_TRY(device_instance_tiles(s, E.count, L.nt, L.tile.as<int>(), d_off, L.tf, L.tl));
    L.w16ok = widest <= 65536;
    if (L.w16ok) VJ_TRY(device_window_offsets(s, L.nt, L.tile.as<int>(), rp, ci, L.win.as<int>(), E.nnz, L.off16));
    VJ_TRY(device_tile_instances(s, L.nt, L.tile.as<int>(), inst_of, L.inst));
    VJ_TRY(device_stage_descs(s, L.nt, L.tile.as<int>(), rp, L.win.as<int>(), L.inst.as<int>(), L.desc));
    VJ_TRY(L.list.alloc(std::max<size_t>(static_cast<size_t>(L.nt), 1) * 4));
    L.hInst.resize(L.nt);
    if (L.nt > 0)
        VJ_HIP(hipMemcpyAsync(L.hInst.data(), L.inst.p, static_cast<size_t>(L.nt) * 4, hipMemcpyDeviceToHost, s));
    std::vector<int> all(L.nt);
    for (int t = 0; t < L.nt; ++t) all[t] = t;
    if (L.nt > 0) VJ_HIP(hipMemcpyAsync(L.list.p, all.data(), static_cast<size_t>(L.nt) * 4, hipMemcpyHostToDevice, s));
    L.nlist = L.nt;
    VJ_HIP(hipStreamSynchronize(s));   // all and hInst are complete
    L.cap = cap;
    return VJ_OK;
}

// Row tiles of the LDS-staged kernels, built on the device once per upload and entry size,
// for both workgroup shapes: for batches (the kernels are autotuner candidates there), and
// for a single LP only when the caller forces them (lanes -5 / -6 / -7): a single LP's tiles
// read windows about as wide as the whole vector, which only a small LP can afford.
int setup_stage(vj_ctx* ctx, FastEngine& E) {
    Problem& P = ctx->prob;
    const vj_params& prm = ctx->params;
    hipStream_t s = ctx->stream;
    const int* rinst = E.multi ? P.row_inst.as<int>() : nullptr;
    const int* cinst = E.multi ? P.col_inst.as<int>() : nullptr;
    const int* roff = E.multi ? P.d_row_off.as<int>() : nullptr;
    const int* coff = E.multi ? P.d_col_off.as<int>() : nullptr;
    const int esz = E.prec == VJ_PREC_FP64 ? 8 : 4;
    for (int sh = 0; sh < 2; ++sh) {
        const int threads = kStageShapeThreads[sh];
        const int cap = threads * kStageLaneBytes / esz;
        if ((E.multi || prm.lanes_A <= -5) && E.stA[sh].cap != cap)
            VJ_TRY(build_stage_layout(s, E, E.stA[sh], threads, E.m, P.rp.as<int>(), P.ci.as<int>(), rinst, roff,
                                      P.d_row_off.as<int>()));
        if ((E.multi || prm.lanes_At <= -5) && E.stAt[sh].cap != cap)
            VJ_TRY(build_stage_layout(s, E, E.stAt[sh], threads, E.n, P.tp.as<int>(), P.ti.as<int>(), cinst, coff,
                                      P.d_col_off.as<int>()));
    }
    return VJ_OK;
}

// The batch launches' tile lists (staged and, for batches, tile-streamed steps): every tile (hs
// null), or the tiles holding a row of an instance active in hs. The caller has drained the
// stream (no kernel reads the lists); an executable graph captured with the old grids is
// released.
int set_launch_lists(vj_ctx* ctx, FastEngine& E, const InstState* hs) {
    hipStream_t s = ctx->stream;
    auto active = [&](int q) { return !hs || hs[q].active != 0; };
    // list the tiles t < count for which live(t) holds, in ascending order
    auto upload = [&](int count, auto live, DevBuf& list, int& n) -> int {
        std::vector<int> h;
        h.reserve(count);
        for (int t = 0; t < count; ++t)
            if (live(t)) h.push_back(t);
        n = static_cast<int>(h.size());
        if (n > 0) VJ_HIP(hipMemcpyAsync(list.p, h.data(), static_cast<size_t>(n) * 4, hipMemcpyHostToDevice, s));
        VJ_HIP(hipStreamSynchronize(s));   // h is released on return
        return VJ_OK;
    };
    // a tile-streamed tile of a batch may hold rows of several instances: listed while one is active
    auto any_active = [&](int first, int last) {
        for (int q = first; q <= last; ++q)
            if (active(q)) return true;
        return false;
    };
    // staged tiles: the running shape's lists; with hs null, every shape's (the autotuner may pick
    // either, and a list pruned by an earlier solve must not carry over)
    for (int sh = 0; sh < 2; ++sh) {
        for (StageLayout* L : {&E.stA[sh], &E.stAt[sh]}) {
            const bool running = L == &E.LA() || L == &E.LAt();
            if (L->cap && (hs == nullptr || running))
                VJ_TRY(upload(L->nt, [&](int t) { return active(L->hInst[t]); }, L->list, L->nlist));
        }
    }
    if (E.multi && E.tiles_built) {
        VJ_TRY(upload(E.ntA, [&](int t) { return any_active(E.hFirstTA[t], E.hLastTA[t]); }, E.listTA, E.nlistTA));
        VJ_TRY(upload(E.ntAt, [&](int t) { return any_active(E.hFirstTAt[t], E.hLastTAt[t]); }, E.listTAt,
                      E.nlistTAt));
    }
    E.pruned = hs != nullptr && ((E.LA().cap && E.LA().nlist < E.LA().nt) || (E.LAt().cap && E.LAt().nlist < E.LAt().nt) ||
                                 (E.multi && E.tiles_built && (E.nlistTA < E.ntA || E.nlistTAt < E.ntAt)));
    if (E.gexec) {
        VJ_HIP(hipGraphExecDestroy(E.gexec));
        E.gexec = nullptr;
    }
    return VJ_OK;
}

// ---- setup: scaling, step sizes, initial centre -------------------------------------
int setup(vj_ctx* ctx, FastEngine& E) {
    Problem& P = ctx->prob;
    const vj_params& prm = ctx->params;
    const int m = E.m, n = E.n, cnt = E.count;
    const int64_t nnz = E.nnz;
    hipStream_t s = ctx->stream;
    const size_t T = E.tsize();
    VJ_TRY(E.av_s.alloc(nnz * 8));
    VJ_TRY(E.tv_s.alloc(nnz * 8));
    VJ_HIP(hipMemcpyAsync(E.av_s.p, P.av.p, nnz * 8, hipMemcpyDeviceToDevice, s));
    VJ_HIP(hipMemcpyAsync(E.tv_s.p, P.tv.p, nnz * 8, hipMemcpyDeviceToDevice, s));
    for (DevBuf* b : {&E.Rsc, &E.rstep, &E.rnorm, &E.ir, &E.pyf, &E.lr_s, &E.ur_s, &E.yhat, &E.yfull, &E.ax_full, &E.wtmp_m})
        VJ_TRY(b->alloc(static_cast<size_t>(m) * 8));
    for (DevBuf* b : {&E.Csc, &E.cstep, &E.cnorm, &E.ic, &E.pxf, &E.c_s, &E.lc_s, &E.uc_s, &E.xhat, &E.xfull,
                      &E.aty_full, &E.vtmp_n, &E.utmp_n})
        VJ_TRY(b->alloc(static_cast<size_t>(n) * 8));
    for (DevBuf* b : {&E.x, &E.xbar, &E.ax, &E.dhat, &E.lcs, &E.ucs}) VJ_TRY(b->alloc(static_cast<size_t>(n) * T));
    for (DevBuf* b : {&E.y, &E.ay, &E.lrs, &E.urs}) VJ_TRY(b->alloc(static_cast<size_t>(m) * T));
    if (E.prec != VJ_PREC_FP64) VJ_TRY(E.yhat_t.alloc(static_cast<size_t>(m) * T));
    if (!E.tiles_built) {   // row tiles of A and A^T for the tile-streamed kernels, built on the device
        VJ_TRY(device_row_tiles(s, m, 1, P.rp.as<int>(), 0, kStreamCap, kBlock, E.tileA, &E.ntA));
        VJ_TRY(device_instance_tiles(s, cnt, E.ntA, E.tileA.as<int>(), P.d_row_off.as<int>(), E.tfA, E.tlA));
        VJ_TRY(device_row_tiles(s, n, 1, P.tp.as<int>(), 0, kStreamCap, kBlock, E.tileAt, &E.ntAt));
        VJ_TRY(device_instance_tiles(s, cnt, E.ntAt, E.tileAt.as<int>(), P.d_col_off.as<int>(), E.tfAt, E.tlAt));
        if (E.multi) {
            VJ_TRY(tile_instance_ranges(s, E.ntA, E.tileA, P.row_off, E.hFirstTA, E.hLastTA));
            VJ_TRY(tile_instance_ranges(s, E.ntAt, E.tileAt, P.col_off, E.hFirstTAt, E.hLastTAt));
            VJ_TRY(E.listTA.alloc(std::max<size_t>(static_cast<size_t>(E.ntA), 1) * 4));
            VJ_TRY(E.listTAt.alloc(std::max<size_t>(static_cast<size_t>(E.ntAt), 1) * 4));
        }
        E.tiles_built = true;
    }
    VJ_TRY(setup_stage(ctx, E));
    const int px_slots = std::max({partial_slots(n, cnt), E.ntAt + cnt + 1, E.stAt[0].nt + cnt + 1, E.stAt[1].nt + cnt + 1});
    const int py_slots = std::max({partial_slots(m, cnt), E.ntA + cnt + 1, E.stA[0].nt + cnt + 1, E.stA[1].nt + cnt + 1});
    VJ_TRY(E.part_x.alloc(static_cast<size_t>(px_slots) * 3 * 8));
    VJ_TRY(E.part_y.alloc(static_cast<size_t>(py_slots) * 3 * 8));
    VJ_TRY(E.part_kr.alloc(static_cast<size_t>(py_slots) * 3 * 8));
    VJ_TRY(E.part_kc.alloc(static_cast<size_t>(px_slots) * 3 * 8));
    VJ_TRY(E.part_s.alloc(static_cast<size_t>(grid_for(std::max(m, n), kBlock) + cnt + 1) * 8));
    VJ_TRY(E.kkt8.alloc(static_cast<size_t>(cnt) * 8 * 8));
    VJ_TRY(E.st.alloc(static_cast<size_t>(cnt) * sizeof(InstState)));
    for (DevBuf* b : {&E.offset_d, &E.norm_b_d, &E.norm_c_d, &E.objfac_d, &E.inst_tmp_a, &E.inst_tmp_b})
        VJ_TRY(b->alloc(static_cast<size_t>(cnt) * 8));
    VJ_HIP(hipMemcpyAsync(E.offset_d.p, P.offset_i.data(), cnt * 8, hipMemcpyHostToDevice, s));
    VJ_HIP(hipMemcpyAsync(E.norm_b_d.p, P.norm_b.data(), cnt * 8, hipMemcpyHostToDevice, s));
    VJ_HIP(hipMemcpyAsync(E.norm_c_d.p, P.norm_c.data(), cnt * 8, hipMemcpyHostToDevice, s));

    const int* rp = P.rp.as<int>();
    const int* ci = P.ci.as<int>();
    const int* tp = P.tp.as<int>();
    const int* ti = P.ti.as<int>();
    double* av = E.av_s.as<double>();
    double* tv = E.tv_s.as<double>();
    const int GA = E.GA, GAt = E.GAt;
    hipLaunchKernelGGL(fill_f64, dim3(grid_for(m, kBlock)), dim3(kBlock), 0, s, m, E.Rsc.as<double>(), 1.0);
    hipLaunchKernelGGL(fill_f64, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, s, n, E.Csc.as<double>(), 1.0);
    VJ_LAUNCHED();

    // Ruiz l_inf equilibration (fixed number of passes), then Pock-Chambolle (alpha = 1).
    const int passes = prm.enable_scaling ? prm.ruiz_iterations : 0;
    for (int pass = 0; pass < passes + (prm.enable_scaling && prm.pock_chambolle ? 1 : 0); ++pass) {
        const int op = pass < passes ? 0 : 1;
        VJ_DISPATCH_G(GA, hipLaunchKernelGGL((row_abs_reduce<GG>), dim3(tiles_for(m, GG)), dim3(kBlock), 0, s, m, rp, av,
                                             op, E.rnorm.as<double>()));
        VJ_DISPATCH_G(GAt, hipLaunchKernelGGL((row_abs_reduce<GG>), dim3(tiles_for(n, GG)), dim3(kBlock), 0, s, n, tp, tv,
                                              op, E.cnorm.as<double>()));
        hipLaunchKernelGGL(make_scale_step, dim3(grid_for(m, kBlock)), dim3(kBlock), 0, s, m, E.rnorm.as<double>(),
                           E.rstep.as<double>(), E.Rsc.as<double>());
        hipLaunchKernelGGL(make_scale_step, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, s, n, E.cnorm.as<double>(),
                           E.cstep.as<double>(), E.Csc.as<double>());
        VJ_DISPATCH_G(GA, hipLaunchKernelGGL((scale_rows<GG>), dim3(tiles_for(m, GG)), dim3(kBlock), 0, s, m, rp, ci, av,
                                             E.rstep.as<double>(), E.cstep.as<double>()));
        VJ_DISPATCH_G(GAt, hipLaunchKernelGGL((scale_rows<GG>), dim3(tiles_for(n, GG)), dim3(kBlock), 0, s, n, tp, ti, tv,
                                              E.cstep.as<double>(), E.rstep.as<double>()));
        VJ_LAUNCHED();
    }
    hipLaunchKernelGGL(scale_cols_data, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, s, n, E.Csc.as<double>(),
                       P.c_can.as<double>(), P.lc.as<double>(), P.uc.as<double>(), E.c_s.as<double>(),
                       E.lc_s.as<double>(), E.uc_s.as<double>());
    hipLaunchKernelGGL(scale_rows_data, dim3(grid_for(m, kBlock)), dim3(kBlock), 0, s, m, E.Rsc.as<double>(),
                       P.lr.as<double>(), P.ur.as<double>(), E.lr_s.as<double>(), E.ur_s.as<double>());
    VJ_LAUNCHED();
    // bound / objective rescaling: cs = 1/(1+|c_s|), bs = 1/(1+|b_s|) per instance, the norms
    // taken without the columns / rows whose equilibration factor reached its bound
    VJ_TRY(seg_norm2(ctx, E, n, 0, E.c_s.as<double>(), nullptr, nullptr, false, E.inst_tmp_a.as<double>(),
                     E.Csc.as<double>()));
    VJ_TRY(seg_norm2(ctx, E, m, 1, nullptr, E.lr_s.as<double>(), E.ur_s.as<double>(), true, E.inst_tmp_b.as<double>(),
                     E.Rsc.as<double>()));
    std::vector<double> cn2(cnt), bn2(cnt), cs(cnt, 1.0), bs(cnt, 1.0), objfac(cnt), omega0(cnt, 1.0);
    VJ_HIP(hipMemcpyAsync(cn2.data(), E.inst_tmp_a.p, cnt * 8, hipMemcpyDeviceToHost, s));
    VJ_HIP(hipMemcpyAsync(bn2.data(), E.inst_tmp_b.p, cnt * 8, hipMemcpyDeviceToHost, s));
    VJ_HIP(hipStreamSynchronize(s));
    for (int q = 0; q < cnt; ++q) {
        const double cnorm = std::sqrt(cn2[q]), bnorm = std::sqrt(bn2[q]);
        if (prm.bound_objective_rescaling) {
            cs[q] = 1.0 / (1.0 + cnorm);
            bs[q] = 1.0 / (1.0 + bnorm);
        }
        objfac[q] = 1.0 / (cs[q] * bs[q]);
        const double cr = cnorm * cs[q], br = bnorm * bs[q];
        omega0[q] = (cr > 1e-10 && br > 1e-10) ? cr / br : 1.0;
    }
    VJ_HIP(hipMemcpyAsync(E.inst_tmp_a.p, cs.data(), cnt * 8, hipMemcpyHostToDevice, s));
    VJ_HIP(hipMemcpyAsync(E.inst_tmp_b.p, bs.data(), cnt * 8, hipMemcpyHostToDevice, s));
    VJ_HIP(hipMemcpyAsync(E.objfac_d.p, objfac.data(), cnt * 8, hipMemcpyHostToDevice, s));
    const int* cinst = E.multi ? P.col_inst.as<int>() : nullptr;
    const int* rinst = E.multi ? P.row_inst.as<int>() : nullptr;
    hipLaunchKernelGGL(rescale_cols, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, s, n, cinst, E.inst_tmp_a.as<double>(),
                       E.inst_tmp_b.as<double>(), E.Csc.as<double>(), E.c_s.as<double>(), E.lc_s.as<double>(),
                       E.uc_s.as<double>(), E.ic.as<double>(), E.pxf.as<double>());
    hipLaunchKernelGGL(rescale_rows, dim3(grid_for(m, kBlock)), dim3(kBlock), 0, s, m, rinst, E.inst_tmp_a.as<double>(),
                       E.inst_tmp_b.as<double>(), E.Rsc.as<double>(), E.lr_s.as<double>(), E.ur_s.as<double>(),
                       E.ir.as<double>(), E.pyf.as<double>());
    VJ_LAUNCHED();

    // Power iteration on A_s^T A_s per instance: lambda ~ |A_s|_2^2.
    double* v = E.vtmp_n.as<double>();
    double* u = E.utmp_n.as<double>();
    double* w = E.wtmp_m.as<double>();
    hipLaunchKernelGGL(init_unit_seg, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, s, n, cinst, P.d_col_off.as<int>(), v);
    VJ_LAUNCHED();
    const int power_iters = prm.power_iterations;
    for (int it = 0; it < power_iters; ++it) {
        VJ_TRY(spmv_f64(ctx, m, rp, ci, av, v, w, GA));
        VJ_TRY(spmv_f64(ctx, n, tp, ti, tv, w, u, GAt));
        VJ_TRY(seg_norm2(ctx, E, n, 0, u, nullptr, nullptr, false, E.inst_tmp_a.as<double>()));
        hipLaunchKernelGGL(normalize_seg, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, s, n, cinst,
                           E.inst_tmp_a.as<double>(), u, v);
        VJ_LAUNCHED();
    }
    std::vector<double> lam2(cnt);
    VJ_HIP(hipMemcpyAsync(lam2.data(), E.inst_tmp_a.p, cnt * 8, hipMemcpyDeviceToHost, s));
    VJ_HIP(hipStreamSynchronize(s));
    E.spectral.assign(cnt, 1.0);
    for (int q = 0; q < cnt; ++q) {
        const double lam = std::sqrt(lam2[q]);  // |A^T A v| with |v| = 1
        E.spectral[q] = lam > 0.0 ? std::sqrt(lam) : 1.0;
    }

    // low-precision matrix copies
    if (E.prec == VJ_PREC_FP32) {
        VJ_TRY(E.av_lo.alloc(nnz * 4));
        VJ_TRY(E.tv_lo.alloc(nnz * 4));
        hipLaunchKernelGGL(convert_f32, dim3(grid_for(nnz, kBlock)), dim3(kBlock), 0, s, nnz, av, E.av_lo.as<float>());
        hipLaunchKernelGGL(convert_f32, dim3(grid_for(nnz, kBlock)), dim3(kBlock), 0, s, nnz, tv, E.tv_lo.as<float>());
        VJ_TRY(E.recA_buf.alloc(nnz * 8));
        VJ_TRY(E.recAt_buf.alloc(nnz * 8));
        hipLaunchKernelGGL(pack_records, dim3(grid_for(nnz, kBlock)), dim3(kBlock), 0, s, nnz, ci, E.av_lo.as<float>(),
                           E.recA_buf.as<vj_i32x2>());
        hipLaunchKernelGGL(pack_records, dim3(grid_for(nnz, kBlock)), dim3(kBlock), 0, s, nnz, ti, E.tv_lo.as<float>(),
                           E.recAt_buf.as<vj_i32x2>());
        VJ_LAUNCHED();
    } else if (E.prec == VJ_PREC_BF16 || E.prec == VJ_PREC_FP16) {
        VJ_TRY(E.av_lo.alloc(nnz * 2));
        VJ_TRY(E.tv_lo.alloc(nnz * 2));
        const bool f16 = E.prec == VJ_PREC_FP16;
        if (f16) {   // the scaled values must fit FP16 (|a| <= 65504); A and A^T hold the same values
            const int parts = 1024;
            DevBuf part;
            VJ_TRY(part.alloc(parts * 8));
            hipLaunchKernelGGL(max_abs_partial, dim3(parts), dim3(kBlock), 0, s, nnz, av, part.as<double>());
            VJ_LAUNCHED();
            std::vector<double> h(parts);
            VJ_HIP(hipMemcpyAsync(h.data(), part.p, parts * 8, hipMemcpyDeviceToHost, s));
            VJ_HIP(hipStreamSynchronize(s));
            const double mx = *std::max_element(h.begin(), h.end());
            if (!(mx <= 65504.0)) {
                set_error("FP16 matrix values: the scaled matrix holds |a| = " + std::to_string(mx) +
                          ", beyond the FP16 range (65504); use fp32 or enable scaling");
                return VJ_ERR_INVALID_ARGUMENT;
            }
            hipLaunchKernelGGL(convert_f16, dim3(grid_for(nnz, kBlock)), dim3(kBlock), 0, s, nnz, av, E.av_lo.as<f16_t>());
            hipLaunchKernelGGL(convert_f16, dim3(grid_for(nnz, kBlock)), dim3(kBlock), 0, s, nnz, tv, E.tv_lo.as<f16_t>());
        } else {
            hipLaunchKernelGGL(convert_bf16, dim3(grid_for(nnz, kBlock)), dim3(kBlock), 0, s, nnz, av,
                               E.av_lo.as<bf16_t>());
            hipLaunchKernelGGL(convert_bf16, dim3(grid_for(nnz, kBlock)), dim3(kBlock), 0, s, nnz, tv,
                               E.tv_lo.as<bf16_t>());
        }
        VJ_LAUNCHED();
        // the staged kernels' 4-byte records, where 16-bit window offsets exist
        auto pack = [&](const DevBuf& off, const DevBuf& val, DevBuf& rec) -> int {
            VJ_TRY(rec.alloc(nnz * 4));
            if (f16)
                hipLaunchKernelGGL(pack_w16<f16_t>, dim3(grid_for(nnz, kBlock)), dim3(kBlock), 0, s, nnz,
                                   off.as<uint16_t>(), val.as<f16_t>(), rec.as<uint32_t>());
            else
                hipLaunchKernelGGL(pack_w16<bf16_t>, dim3(grid_for(nnz, kBlock)), dim3(kBlock), 0, s, nnz,
                                   off.as<uint16_t>(), val.as<bf16_t>(), rec.as<uint32_t>());
            VJ_LAUNCHED();
            return VJ_OK;
        };
        for (int sh = 0; sh < 2; ++sh) {
            if (E.stA[sh].w16ok) VJ_TRY(pack(E.stA[sh].off16, E.av_lo, E.stA[sh].w16rec));
            if (E.stAt[sh].w16ok) VJ_TRY(pack(E.stAt[sh].off16, E.tv_lo, E.stAt[sh].w16rec));
        }
    }

    // state
    std::vector<InstState> hs(cnt);
    for (int q = 0; q < cnt; ++q) {
        InstState& z = hs[q];
        std::memset(&z, 0, sizeof(z));
        z.eta = prm.step_scale / E.spectral[q];
        z.omega = omega0[q];
        z.tau = z.eta / z.omega;
        z.sigma = z.eta * z.omega;
        z.norm_b = P.norm_b[q];
        z.norm_c = P.norm_c[q];
        z.obj_scale = cs[q];
        z.bound_scale = bs[q];
        z.tol = prm.tolerance;
        z.max_iters = prm.max_iterations;
        z.active = 1;
        z.restarted_prev = 1;
        z.r0 = INFINITY;
        z.r_prev = INFINITY;
        z.time_to_target = -1.0;
    }
    VJ_HIP(hipMemcpyAsync(E.st.p, hs.data(), cnt * sizeof(InstState), hipMemcpyHostToDevice, s));

    // initial centre: x0 = clamp(0, lc_s, uc_s), y0 = 0
    hipLaunchKernelGGL(initial_point, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, s, n, E.lc_s.as<double>(),
                       E.uc_s.as<double>(), E.xfull.as<double>());
    hipLaunchKernelGGL(fill_f64, dim3(grid_for(m, kBlock)), dim3(kBlock), 0, s, m, E.yfull.as<double>(), 0.0);
    VJ_LAUNCHED();
    VJ_TRY(launch_kkt(ctx, &E, GA, GAt, E.multi, E.xfull.as<double>(), E.yfull.as<double>(), true, 0,
                      E.part_kr.as<double>(), E.part_kc.as<double>(), E.kkt8.as<double>()));
    // host staging vectors of this function are read by queued copies: finish them here
    VJ_HIP(hipStreamSynchronize(s));
    return VJ_OK;
}

// Column-split copies of the scaled matrix (single LP). The split pattern depends only on
// the sparsity pattern and S, so it is built on the device once per upload and S; the values
// are gathered on the device from the scaled FP64 matrix in the working precision.
int setup_one_split(vj_ctx* ctx, FastEngine& E, int S, int rows, int cols, const int* rp, const int* ci,
                    const double* vals_dev, int& built, DevBuf& bp, DevBuf& idx, DevBuf& perm, DevBuf& val,
                    DevBuf& psum, DevBuf& tiles, DevBuf& tbase, int& ntiles, DevBuf& rec, int& tiles_cap,
                    DevBuf& toff) {
    hipStream_t s = ctx->stream;
    const int64_t nnz = E.nnz;
    if (built != S) {
        VJ_TRY(device_split(s, rows, cols, nnz, rp, ci, S, bp, idx, perm));
        tiles_cap = 0;
        built = S;
    }
    // tiles depend on the LDS capacity in entries and on the tile height of the working
    // precision (both differ between FP64 and the 4-byte iterates, so cap keys the rebuild)
    const int cap = E.prec == VJ_PREC_FP64 ? split_cap<double>() : split_cap<float>();
    const int max_rows = E.prec == VJ_PREC_FP64 ? split_rows<double>() : split_rows<float>();
    if (tiles_cap != cap) {
        const int64_t stride = static_cast<int64_t>(rows) + 1;
        VJ_TRY(device_row_tiles(s, rows, S, bp.as<int>(), stride, cap, max_rows, tiles, &ntiles));
        VJ_TRY(device_tile_bases(s, ntiles, S, tiles.as<int>(), bp.as<int>(), stride, tbase));
        VJ_TRY(device_tile_offsets(s, rows, S, ntiles, tiles.as<int>(), bp.as<int>(), stride, toff));
        tiles_cap = cap;
    }
    const size_t vsz = E.prec == VJ_PREC_FP64 ? 8 : (E.prec == VJ_PREC_FP32 ? 4 : 2);
    VJ_TRY(val.alloc(static_cast<size_t>(nnz) * vsz));
    VJ_TRY(psum.alloc(static_cast<size_t>(S) * static_cast<size_t>(split_pstride(rows)) * E.tsize()));
    const int g = grid_for(nnz, kBlock);
    if (E.prec == VJ_PREC_FP64)
        hipLaunchKernelGGL(gather_split_vals<double>, dim3(g), dim3(kBlock), 0, s, nnz, perm.as<int>(), vals_dev, val.as<double>());
    else if (E.prec == VJ_PREC_FP32) {
        hipLaunchKernelGGL(gather_split_vals<float>, dim3(g), dim3(kBlock), 0, s, nnz, perm.as<int>(), vals_dev, val.as<float>());
        VJ_TRY(rec.alloc(static_cast<size_t>(nnz) * 8));
        hipLaunchKernelGGL(pack_records, dim3(g), dim3(kBlock), 0, s, nnz, idx.as<int>(), val.as<float>(),
                           rec.as<vj_i32x2>());
    }
    else if (E.prec == VJ_PREC_FP16)
        hipLaunchKernelGGL(gather_split_vals<f16_t>, dim3(g), dim3(kBlock), 0, s, nnz, perm.as<int>(), vals_dev, val.as<f16_t>());
    else
        hipLaunchKernelGGL(gather_split_vals<bf16_t>, dim3(g), dim3(kBlock), 0, s, nnz, perm.as<int>(), vals_dev, val.as<bf16_t>());
    VJ_LAUNCHED();
    return VJ_OK;
}

int setup_splits(vj_ctx* ctx, FastEngine& E) {
    Problem& P = ctx->prob;
    const vj_params& prm = ctx->params;
    E.SA = E.SAt = 1;
    if (E.count != 1) return VJ_OK;
    E.SA = prm.col_splits_A ? prm.col_splits_A : choose_splits(E.n);
    E.SAt = prm.col_splits_At ? prm.col_splits_At : choose_splits(E.m);
    if (E.SA > 1)
        VJ_TRY(setup_one_split(ctx, E, E.SA, E.m, E.n, P.rp.as<int>(), P.ci.as<int>(), E.av_s.as<double>(), E.builtSA,
                               E.bpA, E.idxA, E.permA, E.valA, E.psumA, E.tileSA, E.tbaseSA, E.ntSA, E.recSA_buf,
                               E.tcapSA, E.toffSA));
    if (E.SAt > 1)
        VJ_TRY(setup_one_split(ctx, E, E.SAt, E.n, E.m, P.tp.as<int>(), P.ti.as<int>(), E.tv_s.as<double>(),
                               E.builtSAt, E.bpAt, E.idxAt, E.permAt, E.valAt, E.psumAt, E.tileSAt, E.tbaseSAt, E.ntSAt,
                               E.recSAt_buf, E.tcapSAt, E.toffSAt));
    // split partial kernels see ~nnz/(rows*S) entries per row: few lanes per row
    E.GsA = (static_cast<double>(E.nnz) / std::max(1, E.m) / E.SA) > 3.0 ? 4 : 2;
    E.GsAt = (static_cast<double>(E.nnz) / std::max(1, E.n) / E.SAt) > 3.0 ? 4 : 2;
    return VJ_OK;
}

// Sliced-ELL copies for a single LP: the layout is built on the device once per upload for a
// matrix whose padding stays within a quarter of its nonzeros (or whose ELL kernel the
// caller forces); the values are gathered per solve from the scaled FP64 matrix.
constexpr double kEllMaxPadding = 1.25;

// FP32 iterates: a check whose FP64 rel-KKT exceeds this factor times the least rel-KKT of
// the epoch triggers a restart, i.e. an FP64 refinement step.
constexpr double kRefineDrift = 2.0;
// 16-bit values refined at every check: the same rule with a wider factor. Within a long
// epoch their iteration can still drift away from the solution while its working-precision
// residual falls (BF16 on the 1M x 2M LP: from rel-KKT 3.7e-8 up to 8e-2 over 1.2M
// iterations without a restart); their rounding noise moves the rel-KKT by more from check
// to check than FP32's, so the factor is 4 (at 1e-6 it never fires: the same iterations as
// without it). With refinement at restarts only they do not use the rule: their noise then
// exceeds even this factor (BF16: 843 restarts in 150 s).
constexpr double kRefineDrift16 = 4.0;

// Where the FP64 centre moves: 1 = at every check (16-bit matrix values by default, whose
// rounding error otherwise grows with the distance the epoch has moved), 0 = at restarts.
int refine_every_check(const vj_params& p, int prec) {
    if (p.refine_every_check != 0) return p.refine_every_check > 0 ? 1 : 0;
    return (prec == VJ_PREC_BF16 || prec == VJ_PREC_FP16) ? 1 : 0;
}

// The drift-restart factor: restart_drift > 1 as given, < 0 off, 0 the default per precision.
double drift_factor(const vj_params& p, int prec) {
    if (p.restart_drift > 1.0) return p.restart_drift;
    if (p.restart_drift != 0.0) return 0.0;
    if (prec == VJ_PREC_FP32) return kRefineDrift;
    if ((prec == VJ_PREC_BF16 || prec == VJ_PREC_FP16) && refine_every_check(p, prec)) return kRefineDrift16;
    return 0.0;
}

int setup_one_ell(vj_ctx* ctx, FastEngine& E, int rows, const int* rp, const int* ci, const double* vals_dev,
                  bool forced, bool& built, double& pad, int64_t& len, DevBuf& off, DevBuf& width, DevBuf& idx,
                  DevBuf& perm, DevBuf& val, DevBuf& rec) {
    hipStream_t s = ctx->stream;
    if (!built) {   // the layout, on the device; no copy when it pads too much (unless forced)
        VJ_TRY(device_ell(s, rows, E.nnz, rp, ci, kEllMaxPadding, forced, &pad, &len, off, width, idx, perm));
        if (len == 0) return VJ_OK;
        built = true;
    }
    if (len == 0) return VJ_OK;
    const size_t vsz = E.prec == VJ_PREC_FP64 ? 8 : (E.prec == VJ_PREC_FP32 ? 4 : 2);
    VJ_TRY(val.alloc(static_cast<size_t>(len) * vsz));
    const int g = grid_for(len, kBlock);
    if (E.prec == VJ_PREC_FP64)
        hipLaunchKernelGGL(gather_split_vals<double>, dim3(g), dim3(kBlock), 0, s, len, perm.as<int>(), vals_dev, val.as<double>());
    else if (E.prec == VJ_PREC_FP32) {
        hipLaunchKernelGGL(gather_split_vals<float>, dim3(g), dim3(kBlock), 0, s, len, perm.as<int>(), vals_dev, val.as<float>());
        VJ_TRY(rec.alloc(static_cast<size_t>(len) * 8));
        hipLaunchKernelGGL(pack_records, dim3(g), dim3(kBlock), 0, s, len, idx.as<int>(), val.as<float>(),
                           rec.as<vj_i32x2>());
    } else if (E.prec == VJ_PREC_FP16) {
        hipLaunchKernelGGL(gather_split_vals<f16_t>, dim3(g), dim3(kBlock), 0, s, len, perm.as<int>(), vals_dev, val.as<f16_t>());
    } else {
        hipLaunchKernelGGL(gather_split_vals<bf16_t>, dim3(g), dim3(kBlock), 0, s, len, perm.as<int>(), vals_dev, val.as<bf16_t>());
    }
    VJ_LAUNCHED();
    return VJ_OK;
}

int setup_ell(vj_ctx* ctx, FastEngine& E) {
    if (E.count != 1) return VJ_OK;
    Problem& P = ctx->prob;
    const vj_params& prm = ctx->params;
    VJ_TRY(setup_one_ell(ctx, E, E.m, P.rp.as<int>(), P.ci.as<int>(), E.av_s.as<double>(), prm.lanes_A <= -3, E.ellBuiltA,
                         E.ellPadA,
                         E.ellLenA, E.ellOffA, E.ellWA, E.ellIdxA, E.ellPermA, E.ellValA, E.ellRecA));
    VJ_TRY(setup_one_ell(ctx, E, E.n, P.tp.as<int>(), P.ti.as<int>(), E.tv_s.as<double>(), prm.lanes_At <= -3,
                         E.ellBuiltAt, E.ellPadAt,
                         E.ellLenAt, E.ellOffAt, E.ellWAt, E.ellIdxAt, E.ellPermAt, E.ellValAt, E.ellRecAt));
    return VJ_OK;
}

template <class T>
int launch_restart(vj_ctx* ctx, FastEngine& E, int force) {
    Problem& P = ctx->prob;
    hipStream_t s = ctx->stream;
    T* yhat_t = yhat_working<T>(E);
    VJ_TRY(require({E.st.p, E.xfull.p, E.aty_full.p, E.c_s.p, E.lc_s.p, E.uc_s.p, E.xhat.p, E.x.p, E.xbar.p, E.dhat.p,
                    E.ax.p, E.ay.p,
                    E.lcs.p, E.ucs.p, E.yfull.p, E.ax_full.p, E.lr_s.p, E.ur_s.p, E.yhat.p, yhat_t, E.y.p,
                    E.lrs.p, E.urs.p, E.multi ? P.row_inst.p : E.st.p, E.multi ? P.col_inst.p : E.st.p},
                   "restart kernels"));
    if (E.multi) {
        hipLaunchKernelGGL((restart_cols<T, true>), dim3(grid_for(E.n, kBlock)), dim3(kBlock), 0, s, E.n,
                           E.st.as<InstState>(), P.col_inst.as<int>(), force, E.xfull.as<double>(), E.aty_full.as<double>(),
                           E.c_s.as<double>(), E.lc_s.as<double>(), E.uc_s.as<double>(), E.xhat.as<double>(), E.x.as<T>(),
                           E.xbar.as<T>(), E.ax.as<T>(), E.dhat.as<T>(), E.lcs.as<T>(), E.ucs.as<T>(), E.rho);
        hipLaunchKernelGGL((restart_rows<T, true>), dim3(grid_for(E.m, kBlock)), dim3(kBlock), 0, s, E.m,
                           E.st.as<InstState>(), P.row_inst.as<int>(), force, E.yfull.as<double>(), E.ax_full.as<double>(),
                           E.lr_s.as<double>(), E.ur_s.as<double>(), E.yhat.as<double>(), yhat_t, E.y.as<T>(),
                           E.ay.as<T>(), E.lrs.as<T>(), E.urs.as<T>());
    } else {
        hipLaunchKernelGGL((restart_cols<T, false>), dim3(grid_for(E.n, kBlock)), dim3(kBlock), 0, s, E.n,
                           E.st.as<InstState>(), P.col_inst.as<int>(), force, E.xfull.as<double>(), E.aty_full.as<double>(),
                           E.c_s.as<double>(), E.lc_s.as<double>(), E.uc_s.as<double>(), E.xhat.as<double>(), E.x.as<T>(),
                           E.xbar.as<T>(), E.ax.as<T>(), E.dhat.as<T>(), E.lcs.as<T>(), E.ucs.as<T>(), E.rho);
        hipLaunchKernelGGL((restart_rows<T, false>), dim3(grid_for(E.m, kBlock)), dim3(kBlock), 0, s, E.m,
                           E.st.as<InstState>(), P.row_inst.as<int>(), force, E.yfull.as<double>(), E.ax_full.as<double>(),
                           E.lr_s.as<double>(), E.ur_s.as<double>(), E.yhat.as<double>(), yhat_t, E.y.as<T>(),
                           E.ay.as<T>(), E.lrs.as<T>(), E.urs.as<T>());
    }
    VJ_LAUNCHED();
    return VJ_OK;
}

// One block: K iterations, the FP64 check, control and the conditional restart.
template <class V, class T>
int enqueue_block(vj_ctx* ctx, FastEngine& E, const V* av, const V* tv) {
    for (int it = 0; it < E.K; ++it) {
        int flags = 0;
        if (it == 0) flags |= 1;
        if (it == E.K - 1) flags |= 2;
        VJ_TRY(launch_primal<V, T>(ctx, E, tv, E.x.as<T>(), E.xbar.as<T>(), E.y.as<T>(), it, flags));
        VJ_TRY(launch_dual<V, T>(ctx, E, av, E.xbar.as<T>(), E.y.as<T>(), it, flags));
    }
    VJ_TRY(launch_kkt(ctx, &E, E.GA, E.GAt, E.multi, E.xfull.as<double>(), E.yfull.as<double>(), true, 1,
                      E.part_kr.as<double>(), E.part_kc.as<double>(), E.kkt8.as<double>()));
    ControlArgs a;
    a.px = E.part_x.as<double>();
    // rows per workgroup of the kernel writing the partials
    a.Rx = E.SAt > 1 ? kBlock * finish_rows<T>() : (E.ellAt ? kBlock : kBlock / E.GAt);
    a.py = E.part_y.as<double>();
    a.Ry = E.SA > 1 ? kBlock * finish_rows<T>() : (E.ellA ? kBlock : kBlock / E.GA);
    const bool sx = E.SAt == 1 && E.streamAt && !E.ellAt, sy = E.SA == 1 && E.streamA && !E.ellA;
    a.xt_first = E.stageAt ? E.LAt().tf.as<int>() : (sx ? E.tfAt.as<int>() : nullptr);
    a.xt_last = E.stageAt ? E.LAt().tl.as<int>() : (sx ? E.tlAt.as<int>() : nullptr);
    a.yt_first = E.stageA ? E.LA().tf.as<int>() : (sy ? E.tfA.as<int>() : nullptr);
    a.yt_last = E.stageA ? E.LA().tl.as<int>() : (sy ? E.tlA.as<int>() : nullptr);
    a.kkt = E.kkt8.as<double>();
    a.row_off = ctx->prob.d_row_off.as<int>();
    a.col_off = ctx->prob.d_col_off.as<int>();
    a.K = E.K;
    a.restart_sufficient = ctx->params.restart_sufficient;
    a.restart_necessary = ctx->params.restart_necessary;
    a.restart_artificial = ctx->params.restart_artificial;
    a.kkt_drift = drift_factor(ctx->params, E.prec);
    a.pw_theta = ctx->params.primal_weight_smoothing;
    a.pw_ki = ctx->params.primal_weight_integral;
    a.refine_every_check = refine_every_check(ctx->params, E.prec);
    a.hist = E.hist_cap > 0 ? E.hist.as<double>() : nullptr;
    a.hist_cap = E.hist_cap;
    hipLaunchKernelGGL(control_kernel, dim3(E.count), dim3(kBlock), 0, ctx->stream, E.st.as<InstState>(), a);
    VJ_LAUNCHED();
    VJ_TRY(launch_restart<T>(ctx, E, 0));
    return VJ_OK;
}

// Time `reps` launches of each iteration kernel (no partials) on scratch copies of the
// iterates; returns average microseconds.
template <class V, class T>
int time_kernels(vj_ctx* ctx, FastEngine& E, const V* av, const V* tv, int reps, double* us_primal, double* us_dual) {
    hipStream_t s = ctx->stream;
    const size_t Tn = static_cast<size_t>(E.n) * E.tsize(), Tm = static_cast<size_t>(E.m) * E.tsize();
    DevBuf xs, xbs, ys;
    VJ_TRY(xs.alloc(Tn));
    VJ_TRY(xbs.alloc(Tn));
    VJ_TRY(ys.alloc(Tm));
    VJ_HIP(hipMemcpyAsync(xs.p, E.x.p, Tn, hipMemcpyDeviceToDevice, s));
    VJ_HIP(hipMemcpyAsync(xbs.p, E.xbar.p, Tn, hipMemcpyDeviceToDevice, s));
    VJ_HIP(hipMemcpyAsync(ys.p, E.y.p, Tm, hipMemcpyDeviceToDevice, s));
    ScopedEvent se0, se1, se2;
    VJ_HIP(hipEventCreate(&se0.e));
    VJ_HIP(hipEventCreate(&se1.e));
    VJ_HIP(hipEventCreate(&se2.e));
    hipEvent_t e0 = se0.e, e1 = se1.e, e2 = se2.e;
    for (int w = 0; w < 3; ++w) {   // warm-up (clocks, caches)
        VJ_TRY(launch_primal<V, T>(ctx, E, tv, xs.as<T>(), xbs.as<T>(), ys.as<T>(), 1, 0));
        VJ_TRY(launch_dual<V, T>(ctx, E, av, xbs.as<T>(), ys.as<T>(), 1, 0));
    }
    VJ_HIP(hipEventRecord(e0, s));
    for (int r = 0; r < reps; ++r) VJ_TRY(launch_primal<V, T>(ctx, E, tv, xs.as<T>(), xbs.as<T>(), ys.as<T>(), 1, 0));
    VJ_HIP(hipEventRecord(e1, s));
    for (int r = 0; r < reps; ++r) VJ_TRY(launch_dual<V, T>(ctx, E, av, xbs.as<T>(), ys.as<T>(), 1, 0));
    VJ_HIP(hipEventRecord(e2, s));
    VJ_HIP(hipEventSynchronize(e2));
    float ms1 = 0, ms2 = 0;
    VJ_HIP(hipEventElapsedTime(&ms1, e0, e1));
    VJ_HIP(hipEventElapsedTime(&ms2, e1, e2));
    *us_primal = 1000.0 * ms1 / reps;
    *us_dual = 1000.0 * ms2 / reps;
    return VJ_OK;
}

// Autotune, per matrix, the SpMV layout of each fused step on the real data: the
// tile-streamed kernel, a lane-group kernel, or the XCD column split (when the gathered
// vector exceeds the L2 budget) with a few lanes per block segment.

template <class V, class T>
int autotune(vj_ctx* ctx, FastEngine& E, const V* av, const V* tv) {
    const vj_params& prm = ctx->params;
    const double avgA = static_cast<double>(E.nnz) / std::max(1, E.m);
    const double avgAt = static_cast<double>(E.nnz) / std::max(1, E.n);
    const bool tuneA = prm.lanes_A == 0, tuneAt = prm.lanes_At == 0;
    const int SmA = E.SA, SmAt = E.SAt;   // split width prepared by setup_splits
    // forced: the caller's col_splits (0 = free choice, 1 = unsplit only, >1 = split only)
    const bool f32 = E.prec == VJ_PREC_FP32;
    // ell: the sliced-ELL copy exists (single LP, little padding, or forced by lanes -3 / -4)
    // staged candidates per workgroup shape: stage[sh] (its tiles exist and qualify), w16[sh]
    // (its 16-bit offsets exist); stage_threads forces one shape (a forced staged kernel runs
    // the 1,024-thread shape unless stage_threads asks for 256)
    const int forced_shape = prm.stage_threads == 256 ? 1 : 0;
    auto shape_ok = [&](int sh) { return prm.stage_threads == 0 || sh == forced_shape; };
    auto staged = [&](double avg, int sh, bool rec, bool w16) {
        StepCfg c{1, choose_lanes(avg), false, rec, false, true, w16};
        c.shape = sh;
        return c;
    };
    auto base = [&](bool tune, int lanes, double avg, int Sm, int forced, bool ell, const bool (&stage)[2],
                    const bool (&w16)[2]) {
        std::vector<StepCfg> c;
        if (!tune) {
            const int sh = forced_shape;
            if (lanes == -7 && w16[sh]) c.push_back(staged(avg, sh, false, true));
            else if (lanes <= -5) c.push_back(staged(avg, sh, lanes == -6 && f32, false));
            else if (lanes <= -3 && ell) c.push_back({1, choose_lanes(avg), false, lanes == -4 && f32, true});
            else if (lanes < 0) c.push_back({Sm, Sm > 1 ? 0 : choose_lanes(avg), Sm == 1, (lanes == -2 || lanes == -4) && f32, false});
            else c.push_back({Sm, lanes, false, false, false});
            return c;
        }
        // With the tie-break margin below an earlier candidate wins near-ties, so the order is
        // that of the measured preference: the column split first where it exists (the gathered
        // vector exceeds the L2 budget; on the 1M x 2M LP it beat the sliced ELL of A^T by 2%,
        // 77.7 against 79.3 us, too close for the margin alone to separate), FP32 (index, value)
        // records before the separate arrays of the same kernel (faster on every gather-bound
        // step measured, e.g. 77.7 against 82.0 us for the split A x_bar step of that LP), and
        // the sliced ELL, whose copy exists only where it pads little, before the tile-streamed
        // kernel (79.4 against 81.0 us for A^T y unsplit)
        if (Sm > 1) {
            if (f32) c.push_back({Sm, 0, false, true, false});   // tile-streamed split
            c.push_back({Sm, 0, false, false, false});
            for (int g : {1, 2, 4}) c.push_back({Sm, g, false, false, false});
        }
        if (forced <= 1 || Sm == 1) {   // unsplit candidates
            if (ell) {
                if (f32) c.push_back({1, choose_lanes(avg), false, true, true});
                c.push_back({1, choose_lanes(avg), false, false, true});
            }
            if (f32) c.push_back({1, choose_lanes(avg), true, true, false});
            c.push_back({1, choose_lanes(avg), true, false, false});
            // LDS-staged (batches whose tiles read narrow windows), the 1,024-thread shape first
            for (int sh = 0; sh < 2; ++sh) {
                if (!stage[sh] || !shape_ok(sh)) continue;
                if (w16[sh]) c.push_back(staged(avg, sh, false, true));
                if (f32) c.push_back(staged(avg, sh, true, false));
                c.push_back(staged(avg, sh, false, false));
            }
            for (int g : {4, 8, 16, 32}) c.push_back({1, g, false, false, false});
        }
        return c;
    };
    // the staged kernels are candidates where their tiles exist (batches) and the windows they
    // copy to LDS add up to at most kStageMaxWindow times the entries they serve
    constexpr int64_t kStageMaxWindow = 8;
    const int esz = E.prec == VJ_PREC_FP64 ? 8 : 4;
    bool stgA[2], stgAt[2], w16A[2], w16At[2];
    for (int sh = 0; sh < 2; ++sh) {
        const int scap = kStageShapeThreads[sh] * kStageLaneBytes / esz;
        stgA[sh] = E.stA[sh].cap == scap && E.stA[sh].gwin <= kStageMaxWindow * E.nnz;
        stgAt[sh] = E.stAt[sh].cap == scap && E.stAt[sh].gwin <= kStageMaxWindow * E.nnz;
        w16A[sh] = E.stA[sh].w16ok;
        w16At[sh] = E.stAt[sh].w16ok;
    }
    std::vector<StepCfg> cA = base(tuneA, prm.lanes_A, avgA, SmA, prm.col_splits_A, E.ellLenA > 0, stgA, w16A);
    std::vector<StepCfg> cAt = base(tuneAt, prm.lanes_At, avgAt, SmAt, prm.col_splits_At, E.ellLenAt > 0, stgAt, w16At);
    auto apply = [](const StepCfg& c, int& S, int& G, int& Gs, bool& stream, bool& rec, bool& ell, bool& stage,
                    bool& w16, int& shape) {
        S = c.S;
        stream = c.S == 1 && c.stream;
        rec = c.rec;
        ell = c.ell;
        stage = c.stage;
        w16 = c.w16;
        shape = c.shape;
        if (c.S > 1) Gs = c.G;
        else G = std::max(c.G, 2);
    };
    apply(cA[0], E.SA, E.GA, E.GsA, E.streamA, E.recA, E.ellA, E.stageA, E.w16A, E.shA);
    apply(cAt[0], E.SAt, E.GAt, E.GsAt, E.streamAt, E.recAt, E.ellAt, E.stageAt, E.w16At, E.shAt);
    E.K_tuned = 0;
    if (E.nnz < 200000 || (cA.size() == 1 && cAt.size() == 1))
        return VJ_OK;  // launch-bound sizes: the first candidate is as good as timing
    for (const TuneDecision& d : E.tuned)
        if (d.prec == E.prec && d.lanes_A == prm.lanes_A && d.lanes_At == prm.lanes_At &&
            d.col_splits_A == prm.col_splits_A && d.col_splits_At == prm.col_splits_At &&
            d.check_frequency == prm.check_frequency && d.stage_threads == prm.stage_threads) {
            apply(d.a, E.SA, E.GA, E.GsA, E.streamA, E.recA, E.ellA, E.stageA, E.w16A, E.shA);
            apply(d.at, E.SAt, E.GAt, E.GsAt, E.streamAt, E.recAt, E.ellAt, E.stageAt, E.w16At, E.shAt);
            E.GA = d.GA;
            E.GAt = d.GAt;
            E.K_tuned = d.K_tuned;
            return VJ_OK;
        }
    // A later candidate replaces the best so far only when it is faster by more than this
    // fraction: near-ties keep the earlier candidate instead of following timing noise.
    constexpr double kTuneMargin = 0.03;
    constexpr int kTuneReps = 10;   // timed launches per candidate and step
    const size_t rounds = std::max(cA.size(), cAt.size());
    double bestA = 1e30, bestAt = 1e30;
    StepCfg wA = cA[0], wAt = cAt[0];
    {   // one untimed pass first: the first candidate is otherwise timed on colder clocks and
        // caches than the rest, which biases the comparison against it
        double up = 0, ud = 0;
        apply(cA[0], E.SA, E.GA, E.GsA, E.streamA, E.recA, E.ellA, E.stageA, E.w16A, E.shA);
        apply(cAt[0], E.SAt, E.GAt, E.GsAt, E.streamAt, E.recAt, E.ellAt, E.stageAt, E.w16At, E.shAt);
        VJ_TRY((time_kernels<V, T>(ctx, E, av, tv, kTuneReps, &up, &ud)));
    }
    for (size_t r = 0; r < rounds; ++r) {
        const StepCfg& a = cA[std::min(r, cA.size() - 1)];
        const StepCfg& b = cAt[std::min(r, cAt.size() - 1)];
        apply(a, E.SA, E.GA, E.GsA, E.streamA, E.recA, E.ellA, E.stageA, E.w16A, E.shA);
        apply(b, E.SAt, E.GAt, E.GsAt, E.streamAt, E.recAt, E.ellAt, E.stageAt, E.w16At, E.shAt);
        double up = 0, ud = 0;
        VJ_TRY((time_kernels<V, T>(ctx, E, av, tv, kTuneReps, &up, &ud)));
        if (prm.verbose >= 2)
            std::fprintf(stderr, "[autotune] A: S=%d G=%d stream=%d rec=%d ell=%d stage=%d w16=%d %.1f us | A^T: S=%d "
                         "G=%d stream=%d rec=%d ell=%d stage=%d w16=%d %.1f us\n",
                         a.S, a.G, a.stream ? 1 : 0, a.rec ? 1 : 0, a.ell ? 1 : 0,
                         a.stage ? kStageShapeThreads[a.shape] : 0, a.w16 ? 1 : 0, ud, b.S, b.G, b.stream ? 1 : 0,
                         b.rec ? 1 : 0, b.ell ? 1 : 0, b.stage ? kStageShapeThreads[b.shape] : 0, b.w16 ? 1 : 0, up);
        if (r == 0 || (r < cA.size() && ud < (1.0 - kTuneMargin) * bestA)) { bestA = ud; wA = a; }
        if (r == 0 || (r < cAt.size() && up < (1.0 - kTuneMargin) * bestAt)) { bestAt = up; wAt = b; }
    }
    apply(wA, E.SA, E.GA, E.GsA, E.streamA, E.recA, E.ellA, E.stageA, E.w16A, E.shA);
    apply(wAt, E.SAt, E.GAt, E.GsAt, E.streamAt, E.recAt, E.ellAt, E.stageAt, E.w16At, E.shAt);
    // The FP64 check passes run lane-group kernels with E.GA / E.GAt lanes per row. A split
    // winner leaves them at the last candidate timed: give them the width that suits the rows.
    if (E.SA > 1 || E.ellA) E.GA = std::max(choose_lanes(avgA), 2);
    if (E.SAt > 1 || E.ellAt) E.GAt = std::max(choose_lanes(avgAt), 2);
    // Iterations per check: the smallest power of two in [32, 256] at which the check (two
    // FP64 KKT passes, their reduction, control and restart kernels) costs at most 2% of the
    // iterations between checks. The KKT passes are timed on the current iterates; their
    // outputs are recomputed by the first real check. BF16 values refined at every check: the
    // check is also the refinement step, whose frequency sets how far the 8-bit values'
    // error grows, so it may cost up to 8%, in [64, 256] (on the 1M x 2M LP K = 64 took
    // 131,136 iterations against 297,472 at K = 256; on the batches 2-7% less time). FP16
    // keeps the 2% rule: at 8% its staircase batch took 3% longer (2.20 s against 2.14 s).
    if (prm.check_frequency == 0) {
        ScopedEvent se0, se1;
        VJ_HIP(hipEventCreate(&se0.e));
        VJ_HIP(hipEventCreate(&se1.e));
        hipEvent_t e0 = se0.e, e1 = se1.e;
        VJ_HIP(hipEventRecord(e0, ctx->stream));
        for (int r = 0; r < 3; ++r)
            VJ_TRY(launch_kkt(ctx, &E, E.GA, E.GAt, E.multi, E.xfull.as<double>(), E.yfull.as<double>(), true, 1,
                              E.part_kr.as<double>(), E.part_kc.as<double>(), E.kkt8.as<double>()));
        VJ_HIP(hipEventRecord(e1, ctx->stream));
        VJ_HIP(hipEventSynchronize(e1));
        float ms = 0.0f;
        VJ_HIP(hipEventElapsedTime(&ms, e0, e1));
        const double t_check = 1000.0 * ms / 3 + 20.0;   // + control and restart kernels
        const double t_iter = bestA + bestAt;
        const bool bf16_refined = E.prec == VJ_PREC_BF16 && refine_every_check(prm, E.prec);
        const double allowance = bf16_refined ? 0.08 : 0.02;
        int K = bf16_refined ? 64 : 32;
        while (K < 256 && t_check / K > allowance * t_iter) K *= 2;
        E.K_tuned = K;
    }
    E.tuned.push_back(TuneDecision{E.prec, prm.lanes_A, prm.lanes_At, prm.col_splits_A, prm.col_splits_At,
                                   prm.check_frequency, prm.stage_threads, wA, wAt, E.GA, E.GAt, E.K_tuned});
    return VJ_OK;
}

// Capture one check block as a hipGraph and instantiate it (E.gexec).
template <class V, class T>
int capture_block(vj_ctx* ctx, FastEngine& E, const V* av, const V* tv) {
    hipStream_t s = ctx->stream;
    hipGraph_t g = nullptr;
    VJ_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    const int rc = enqueue_block<V, T>(ctx, E, av, tv);
    hipError_t ec = hipStreamEndCapture(s, &g);
    if (rc != VJ_OK) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    VJ_HIP(ec);
    const hipError_t ei = hipGraphInstantiate(&E.gexec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);   // the executable graph keeps what it needs
    VJ_HIP(ei);
    return VJ_OK;
}

// Batch launches (LDS-staged and tile-streamed steps): when at most half of the tiles a launch
// lists hold rows of instances still active (by the state polled one block behind), drain the
// stream and list only the tiles of the instances active now; the block graph is captured
// again with the new grids.
template <class V, class T>
int prune_launch_lists(vj_ctx* ctx, FastEngine& E, const InstState* polled, const InstState* latest, bool graph,
                      const V* av, const V* tv) {
    if (E.count < 2) return VJ_OK;
    auto live = [&](const std::vector<int>& inst) {
        int n = 0;
        for (int q : inst) n += polled[q].active != 0;
        return n;
    };
    auto live_ranges = [&](const std::vector<int>& first, const std::vector<int>& last) {
        int n = 0;
        for (size_t t = 0; t < first.size(); ++t) {
            bool l = false;
            for (int q = first[t]; !l && q <= last[t]; ++q) l = polled[q].active != 0;
            n += l;
        }
        return n;
    };
    const bool sa = E.SA == 1 && E.streamA && !E.ellA && !E.stageA;   // tile-streamed steps of the batch
    const bool sat = E.SAt == 1 && E.streamAt && !E.ellAt && !E.stageAt;
    // rebuild when at most kPruneLive of the listed tiles are live (0.75 measured the same on
    // both batches: the finished tiles' workgroups exit after two loads)
    constexpr double kPruneLive = 0.5;
    const bool a = (E.stageA && live(E.LA().hInst) <= kPruneLive * E.LA().nlist) ||
                   (sa && live_ranges(E.hFirstTA, E.hLastTA) <= kPruneLive * E.nlistTA);
    const bool at = (E.stageAt && live(E.LAt().hInst) <= kPruneLive * E.LAt().nlist) ||
                    (sat && live_ranges(E.hFirstTAt, E.hLastTAt) <= kPruneLive * E.nlistTAt);
    if (!a && !at) return VJ_OK;
    VJ_HIP(hipStreamSynchronize(ctx->stream));   // `latest` now holds the state after the last block
    VJ_TRY(set_launch_lists(ctx, E, latest));
    if (graph) VJ_TRY((capture_block<V, T>(ctx, E, av, tv)));
    return VJ_OK;
}

template <class V, class T>
int run_typed(vj_ctx* ctx, FastEngine& E, vj_result* out, clk::time_point t_start) {
    hipStream_t s = ctx->stream;
    const vj_params& prm = ctx->params;
    const V* av = nullptr;
    const V* tv = nullptr;
    if constexpr (sizeof(V) == 8) {
        av = reinterpret_cast<const V*>(E.av_s.p);
        tv = reinterpret_cast<const V*>(E.tv_s.p);
    } else {
        av = reinterpret_cast<const V*>(E.av_lo.p);
        tv = reinterpret_cast<const V*>(E.tv_lo.p);
    }
    const auto t_layout = clk::now();
    VJ_TRY(setup_splits(ctx, E));
    VJ_TRY(setup_ell(ctx, E));
    VJ_HIP(hipStreamSynchronize(s));
    const double layout_time = secs(t_layout, clk::now());
    if (E.stA[0].cap || E.stAt[0].cap || E.multi) VJ_TRY(set_launch_lists(ctx, E, nullptr));   // every tile
    VJ_TRY(launch_restart<T>(ctx, E, 1));  // initial centring
    VJ_TRY((autotune<V, T>(ctx, E, av, tv)));
    // after autotuning the scratch runs left the real iterates untouched; recentre anyway
    VJ_TRY(launch_restart<T>(ctx, E, 1));
    E.K = prm.check_frequency > 0 ? prm.check_frequency : (E.K_tuned > 0 ? E.K_tuned : 64);
    VJ_HIP(hipStreamSynchronize(s));
    const auto t_iter0 = clk::now();

    const int cnt = E.count;
    if (E.h_state_cap < cnt) {
        for (auto*& p : E.h_state)
            if (p) {
                VJ_HIP(hipHostFree(p));
                p = nullptr;
            }
        E.h_state_cap = 0;
        for (auto*& p : E.h_state) VJ_HIP(hipHostMalloc(reinterpret_cast<void**>(&p), cnt * sizeof(InstState), 0));
        E.h_state_cap = cnt;
    }
    bool use_graph = prm.use_graph != 0;
    if (use_graph) VJ_TRY((capture_block<V, T>(ctx, E, av, tv)));
    ScopedEvent sev[2], sev0;
    VJ_HIP(hipEventCreate(&sev[0].e));
    VJ_HIP(hipEventCreate(&sev[1].e));
    VJ_HIP(hipEventCreate(&sev0.e));
    const hipEvent_t ev[2] = {sev[0].e, sev[1].e};
    const hipEvent_t ev0 = sev0.e;
    VJ_HIP(hipEventRecord(ev0, s));
    std::vector<double> conv_time(cnt, -1.0);
    int64_t transfers = 0;
    bool all_done = false, timed_out = false;
    auto t_report = clk::now();   // verbose >= 1: a progress line every 10 s of a long solve
    for (int64_t b = 0;; ++b) {
        if (use_graph) VJ_HIP(hipGraphLaunch(E.gexec, s));
        else VJ_TRY((enqueue_block<V, T>(ctx, E, av, tv)));
        VJ_HIP(hipMemcpyAsync(E.h_state[b & 1], E.st.p, cnt * sizeof(InstState), hipMemcpyDeviceToHost, s));
        VJ_HIP(hipEventRecord(ev[b & 1], s));
        ++transfers;
        if (b >= 1) {
            VJ_HIP(hipEventSynchronize(ev[(b - 1) & 1]));
            const InstState* hs = E.h_state[(b - 1) & 1];
            float ms = 0.0f;
            VJ_HIP(hipEventElapsedTime(&ms, ev0, ev[(b - 1) & 1]));
            const double t_now = 1e-3 * ms;  // device time at the end of the observed block
            int active = 0;
            for (int q = 0; q < cnt; ++q) {
                if (hs[q].active) ++active;
                if (hs[q].converged && conv_time[q] < 0.0) conv_time[q] = t_now;
            }
            if (active == 0) { all_done = true; break; }
            if (secs(t_start, clk::now()) > prm.time_limit) { timed_out = true; break; }
            VJ_TRY((prune_launch_lists<V, T>(ctx, E, hs, E.h_state[b & 1], use_graph, av, tv)));
            if (prm.verbose >= 1 && secs(t_report, clk::now()) >= 10.0) {
                t_report = clk::now();
                int64_t it = 0;
                double worst = 0.0;
                for (int q = 0; q < cnt; ++q) {
                    it = std::max<int64_t>(it, hs[q].iters);
                    if (hs[q].active) worst = std::max(worst, hs[q].rel_kkt);
                }
                std::fprintf(stderr, "[vj] %.1f s: %lld iterations, worst live rel-KKT %.3e, %d/%d live, restarts %d\n",
                             t_now, static_cast<long long>(it), worst, active, cnt, hs[0].restarts);
                std::fflush(stderr);
            }
        }
    }
    VJ_HIP(hipStreamSynchronize(s));
    const auto t_iter1 = clk::now();
    (void)all_done;

    // final state and solution (original space)
    std::vector<InstState> hs(cnt);
    VJ_HIP(hipMemcpy(hs.data(), E.st.p, cnt * sizeof(InstState), hipMemcpyDeviceToHost));
    Problem& P = ctx->prob;
    VJ_TRY(ctx->x_sol.alloc(static_cast<size_t>(E.n) * 8));
    VJ_TRY(ctx->y_sol.alloc(static_cast<size_t>(E.m) * 8));
    hipLaunchKernelGGL(unscale_out, dim3(grid_for(E.n, kBlock)), dim3(kBlock), 0, s, E.n, E.xfull.as<double>(),
                       E.pxf.as<double>(), ctx->x_sol.as<double>());
    hipLaunchKernelGGL(unscale_out, dim3(grid_for(E.m, kBlock)), dim3(kBlock), 0, s, E.m, E.yfull.as<double>(),
                       E.pyf.as<double>(), ctx->y_sol.as<double>());
    VJ_LAUNCHED();
    ctx->have_solution = true;
    std::vector<double> audit(static_cast<size_t>(cnt) * 8);
    VJ_TRY(audit_device(ctx, ctx->x_sol.as<double>(), ctx->y_sol.as<double>(), audit.data()));
    if (E.hist_cap > 0) {
        E.h_hist.resize(static_cast<size_t>(cnt) * E.hist_cap);
        VJ_HIP(hipMemcpy(E.h_hist.data(), E.hist.p, E.h_hist.size() * 8, hipMemcpyDeviceToHost));
    }
    const auto t_end = clk::now();
    for (int q = 0; q < cnt; ++q) {
        vj_result& r = out[q];
        std::memset(&r, 0, sizeof(r));
        const InstState& z = hs[q];
        r.converged = z.converged;
        r.iterations = z.iters;
        r.restart_count = z.restarts;
        r.rel_kkt = audit[8 * q + 0];
        r.rel_primal = audit[8 * q + 1];
        r.rel_dual = audit[8 * q + 2];
        r.rel_gap = audit[8 * q + 3];
        r.primal_objective = audit[8 * q + 4];
        r.dual_objective = audit[8 * q + 5];
        r.spectral_norm = E.spectral[q];
        r.tau = z.tau;
        r.sigma = z.sigma;
        r.primal_weight = z.omega;
        r.primal_residual = z.rel_p;
        r.dual_residual = z.rel_d;
        r.duality_gap = z.rel_gap;
        r.iteration_time = secs(t_iter0, t_iter1);
        r.total_time = secs(t_start, t_end);
        r.time_to_target = conv_time[q];
        r.host_device_transfers = transfers;
        r.refinements = E.prec == VJ_PREC_FP64 ? 0 : z.restarts;
        // 0: tile-streamed kernel, -3 / -4: sliced ELL, -5 / -6: LDS-staged (separate arrays /
        // FP32 records), -7: LDS-staged with 16-bit window offsets
        r.lanes_A = E.stageA ? (E.w16A ? -7 : (E.recA ? -6 : -5))
                             : (E.ellA ? (E.recA ? -4 : -3) : (E.SA > 1 ? E.GsA : (E.streamA ? 0 : E.GA)));
        r.lanes_At = E.stageAt ? (E.w16At ? -7 : (E.recAt ? -6 : -5))
                               : (E.ellAt ? (E.recAt ? -4 : -3) : (E.SAt > 1 ? E.GsAt : (E.streamAt ? 0 : E.GAt)));
        r.col_splits_A = E.SA;
        r.col_splits_At = E.SAt;
        r.check_interval = E.K;
        r.checks = z.checks;
        r.layout_time = layout_time;
        if (z.nan_flag) r.status = VJ_STATUS_NUMERICAL_ERROR;
        else if (z.converged) r.status = VJ_STATUS_OPTIMAL;
        else if (timed_out) r.status = VJ_STATUS_TIME_LIMIT;
        else r.status = VJ_STATUS_ITERATION_LIMIT;
        (void)P;
    }
    return VJ_OK;
}

}  // namespace

int audit_device(vj_ctx* ctx, const double* d_x, const double* d_y, double* out8) {
    Problem& P = ctx->prob;
    const int cnt = P.count;
    DevBuf pr, pc, o8, offs, nb, nc;
    VJ_TRY(pr.alloc(static_cast<size_t>(partial_slots(P.m, cnt)) * 3 * 8));
    VJ_TRY(pc.alloc(static_cast<size_t>(partial_slots(P.n, cnt)) * 3 * 8));
    VJ_TRY(o8.alloc(static_cast<size_t>(cnt) * 64));
    VJ_TRY(offs.alloc(static_cast<size_t>(cnt) * 8));
    VJ_TRY(nb.alloc(static_cast<size_t>(cnt) * 8));
    VJ_TRY(nc.alloc(static_cast<size_t>(cnt) * 8));
    hipStream_t s = ctx->stream;
    VJ_HIP(hipMemcpyAsync(offs.p, P.offset_i.data(), cnt * 8, hipMemcpyHostToDevice, s));
    VJ_HIP(hipMemcpyAsync(nb.p, P.norm_b.data(), cnt * 8, hipMemcpyHostToDevice, s));
    VJ_HIP(hipMemcpyAsync(nc.p, P.norm_c.data(), cnt * 8, hipMemcpyHostToDevice, s));
    const int GA = choose_lanes(static_cast<double>(P.nnz) / std::max(1, P.m));
    const int GAt = choose_lanes(static_cast<double>(P.nnz) / std::max(1, P.n));
    // the audit reads the original data only; a temporary engine view carries the norms
    FastEngine view;
    view.norm_b_d.p = nb.p;
    view.norm_c_d.p = nc.p;
    view.offset_d.p = offs.p;
    int rc = launch_kkt(ctx, &view, GA, GAt, cnt > 1, d_x, d_y, false, 0, pr.as<double>(), pc.as<double>(),
                        o8.as<double>());
    view.norm_b_d.p = nullptr;  // not owned
    view.norm_c_d.p = nullptr;
    view.offset_d.p = nullptr;
    if (rc != VJ_OK) return rc;
    VJ_HIP(hipMemcpyAsync(out8, o8.p, static_cast<size_t>(cnt) * 64, hipMemcpyDeviceToHost, s));
    VJ_HIP(hipStreamSynchronize(s));
    return VJ_OK;
}

int fast_solve(vj_ctx* ctx, vj_result* out) {
    const auto t0 = clk::now();
    Problem& P = ctx->prob;
    if (!ctx->fast) ctx->fast.reset(new FastEngine());
    FastEngine& E = *ctx->fast;
    // The previous solve's executable graph holds the device addresses of this engine's
    // buffers, which the setup below frees and re-allocates (DevBuf::alloc releases first):
    // drain the stream and destroy that graph before any buffer changes hands.
    VJ_HIP(hipStreamSynchronize(ctx->stream));
    if (E.gexec) {
        VJ_HIP(hipGraphExecDestroy(E.gexec));
        E.gexec = nullptr;
    }
    E.prec = ctx->params.precision;
    E.count = P.count;
    E.m = P.m;
    E.n = P.n;
    E.nnz = P.nnz;
    E.multi = P.count > 1;
    E.rho = ctx->params.reflection;
    E.GA = ctx->params.lanes_A > 0 ? ctx->params.lanes_A : choose_lanes(static_cast<double>(P.nnz) / std::max(1, P.m));
    E.GAt = ctx->params.lanes_At > 0 ? ctx->params.lanes_At : choose_lanes(static_cast<double>(P.nnz) / std::max(1, P.n));
    E.streamA = ctx->params.lanes_A == 0;
    E.streamAt = ctx->params.lanes_At == 0;
    E.hist_cap = P.count == 1 ? 1 << 14 : 0;
    if (E.hist_cap) VJ_TRY(E.hist.alloc(static_cast<size_t>(E.hist_cap) * 8));
    VJ_TRY(setup(ctx, E));
    VJ_HIP(hipStreamSynchronize(ctx->stream));
    const double setup_time = secs(t0, clk::now());
    int rc;
    switch (E.prec) {
        case VJ_PREC_FP64: rc = run_typed<double, double>(ctx, E, out, t0); break;
        case VJ_PREC_FP32: rc = run_typed<float, float>(ctx, E, out, t0); break;
        case VJ_PREC_BF16: rc = run_typed<bf16_t, float>(ctx, E, out, t0); break;
        case VJ_PREC_FP16: rc = run_typed<f16_t, float>(ctx, E, out, t0); break;
        default: set_error("unknown precision"); return VJ_ERR_INVALID_ARGUMENT;
    }
    if (rc != VJ_OK) return rc;
    for (int q = 0; q < P.count; ++q) out[q].setup_time = setup_time;
    return VJ_OK;
}

namespace {

// Enqueue `count` check blocks of the last solve: its executable graph when one was
// captured, else the same launches on the stream.
template <class V, class T>
int enqueue_blocks(vj_ctx* ctx, FastEngine& E, int count) {
    const V* av = reinterpret_cast<const V*>(sizeof(V) == 8 ? E.av_s.p : E.av_lo.p);
    const V* tv = reinterpret_cast<const V*>(sizeof(V) == 8 ? E.tv_s.p : E.tv_lo.p);
    for (int b = 0; b < count; ++b) {
        if (E.gexec) VJ_HIP(hipGraphLaunch(E.gexec, ctx->stream));
        else VJ_TRY((enqueue_block<V, T>(ctx, E, av, tv)));
    }
    return VJ_OK;
}

int enqueue_blocks_any(vj_ctx* ctx, FastEngine& E, int count) {
    switch (E.prec) {
        case VJ_PREC_FP64: return enqueue_blocks<double, double>(ctx, E, count);
        case VJ_PREC_FP32: return enqueue_blocks<float, float>(ctx, E, count);
        case VJ_PREC_FP16: return enqueue_blocks<f16_t, float>(ctx, E, count);
        default: return enqueue_blocks<bf16_t, float>(ctx, E, count);
    }
}

}  // namespace

int fast_benchmark_blocks(vj_ctx* ctx, int warmup, int steps, double* ms_per_block, int* iters_per_block) {
    if (!ctx->fast || !ctx->have_solution) {
        set_error("no restarted-mode solve has run on this context");
        return VJ_ERR_NOT_READY;
    }
    FastEngine& E = *ctx->fast;
    hipStream_t s = ctx->stream;
    const size_t sb = static_cast<size_t>(E.count) * sizeof(InstState);
    std::vector<InstState> saved(E.count), live;
    VJ_HIP(hipMemcpyAsync(saved.data(), E.st.p, sb, hipMemcpyDeviceToHost, s));
    VJ_HIP(hipStreamSynchronize(s));
    // every instance iterates below: staged launches list every tile again (the graph of the
    // solve, captured with pruned grids, is released and the blocks are launched directly)
    if (E.pruned) VJ_TRY(set_launch_lists(ctx, E, nullptr));
    // every instance iterates through all blocks: none may converge or reach its limit
    live = saved;
    for (auto& z : live) {
        z.active = 1;
        z.converged = 0;
        z.tol = 0.0;
        z.max_iters = INT64_MAX / 4;
    }
    VJ_HIP(hipMemcpyAsync(E.st.p, live.data(), sb, hipMemcpyHostToDevice, s));
    ScopedEvent se0, se1;
    VJ_HIP(hipEventCreate(&se0.e));
    VJ_HIP(hipEventCreate(&se1.e));
    int rc = enqueue_blocks_any(ctx, E, warmup);
    if (rc == VJ_OK) {
        VJ_HIP(hipEventRecord(se0.e, s));
        rc = enqueue_blocks_any(ctx, E, steps);
    }
    VJ_HIP(hipEventRecord(se1.e, s));
    VJ_HIP(hipEventSynchronize(se1.e));
    // the returned point was saved before; the final state of the solve goes back
    VJ_HIP(hipMemcpyAsync(E.st.p, saved.data(), sb, hipMemcpyHostToDevice, s));
    VJ_HIP(hipStreamSynchronize(s));
    if (rc != VJ_OK) return rc;
    float ms = 0.0f;
    VJ_HIP(hipEventElapsedTime(&ms, se0.e, se1.e));
    *ms_per_block = static_cast<double>(ms) / steps;
    *iters_per_block = E.K;
    return VJ_OK;
}

int fast_scaling(vj_ctx* ctx, double* row_scale, double* col_scale) {
    if (!ctx->fast || !ctx->fast->Rsc.p || !ctx->fast->Csc.p) {
        set_error("no restarted-mode solve has run on this context");
        return VJ_ERR_NOT_READY;
    }
    FastEngine& E = *ctx->fast;
    if (row_scale)
        VJ_HIP(hipMemcpyAsync(row_scale, E.Rsc.p, static_cast<size_t>(E.m) * 8, hipMemcpyDeviceToHost, ctx->stream));
    if (col_scale)
        VJ_HIP(hipMemcpyAsync(col_scale, E.Csc.p, static_cast<size_t>(E.n) * 8, hipMemcpyDeviceToHost, ctx->stream));
    VJ_HIP(hipStreamSynchronize(ctx->stream));
    return VJ_OK;
}

int fast_benchmark(vj_ctx* ctx, int reps, vj_kernel_stat* out, int cap, int* written) {
    *written = 0;
    if (!ctx->fast) {
        set_error("no restarted-mode solve has run on this context");
        return VJ_ERR_NOT_READY;
    }
    FastEngine& E = *ctx->fast;
    double up = 0, ud = 0;
    int rc;
    // after a solve every instance is inactive and the kernels would exit at once: time
    // them with all instances marked active, then put the final state back
    std::vector<InstState> saved(E.count), active(E.count);
    VJ_HIP(hipMemcpy(saved.data(), E.st.p, E.count * sizeof(InstState), hipMemcpyDeviceToHost));
    if (E.pruned) {   // staged launches list every tile again
        VJ_HIP(hipStreamSynchronize(ctx->stream));
        VJ_TRY(set_launch_lists(ctx, E, nullptr));
    }
    active = saved;
    for (auto& z : active) z.active = 1;
    VJ_HIP(hipMemcpy(E.st.p, active.data(), E.count * sizeof(InstState), hipMemcpyHostToDevice));
    switch (E.prec) {
        case VJ_PREC_FP64:
            rc = time_kernels<double, double>(ctx, E, E.av_s.as<double>(), E.tv_s.as<double>(), reps, &up, &ud);
            break;
        case VJ_PREC_FP32:
            rc = time_kernels<float, float>(ctx, E, E.av_lo.as<float>(), E.tv_lo.as<float>(), reps, &up, &ud);
            break;
        default:
            if (E.prec == VJ_PREC_FP16)
                rc = time_kernels<f16_t, float>(ctx, E, E.av_lo.as<f16_t>(), E.tv_lo.as<f16_t>(), reps, &up, &ud);
            else
                rc = time_kernels<bf16_t, float>(ctx, E, E.av_lo.as<bf16_t>(), E.tv_lo.as<bf16_t>(), reps, &up, &ud);
            break;
    }
    VJ_HIP(hipMemcpy(E.st.p, saved.data(), E.count * sizeof(InstState), hipMemcpyHostToDevice));
    if (rc != VJ_OK) return rc;
    const double sv = E.prec == VJ_PREC_FP64 ? 8 : (E.prec == VJ_PREC_FP32 ? 4 : 2);
    const double sx = E.prec == VJ_PREC_FP64 ? 8 : 4;
    const double m = E.m, n = E.n, nnz = static_cast<double>(E.nnz);
    // algorithmic bytes: matrix once, pointers once, gathered vector once, per-row streams
    // primal_step rows of A^T: reads y (gather, m), x (or xbar), dhat, lcs, ucs (4n); writes
    // xbar, and x unless the reflection is 1 (x is then recovered from xbar)
    const double b_primal = nnz * (sv + 4) + (n + 1) * 4 + m * sx + (E.rho == 1.0 ? 5 : 6) * n * sx;
    // dual_step rows of A: reads xbar (gather, n), y, yhat, lrs, urs (4m); writes y (m)
    const double b_dual = nnz * (sv + 4) + (m + 1) * 4 + n * sx + 5 * m * sx;
    // names carry the kernel variant the autotuner chose, as the profiler lists it
    const char* pt = E.prec == VJ_PREC_FP64 ? "double"
                   : (E.prec == VJ_PREC_FP32 ? "float" : (E.prec == VJ_PREC_FP16 ? "f16" : "bf16"));
    char names[2][48];
    const char* ra = E.recA ? ",rec" : "";
    const char* rt = E.recAt ? ",rec" : "";
    if (E.SAt > 1 && E.GsAt == 0) std::snprintf(names[0], 48, "split_stream_partials+primal_finish<%s%s>", pt, rt);
    else if (E.SAt > 1) std::snprintf(names[0], 48, "split_partials+primal_finish<%s,G=%d>", pt, E.GsAt);
    else if (E.ellAt) std::snprintf(names[0], 48, "ell_primal_step<%s%s>", pt, rt);
    else if (E.stageAt)
        std::snprintf(names[0], 48, "staged_primal_step<%s%s,%d>", pt, E.w16At ? ",w16" : rt, E.LAt().threads);
    else if (E.streamAt) std::snprintf(names[0], 48, "stream_primal_step<%s%s>", pt, rt);
    else std::snprintf(names[0], 48, "primal_step<%s,G=%d>", pt, E.GAt);
    if (E.SA > 1 && E.GsA == 0) std::snprintf(names[1], 48, "split_stream_partials+dual_finish<%s%s>", pt, ra);
    else if (E.SA > 1) std::snprintf(names[1], 48, "split_partials+dual_finish<%s,G=%d>", pt, E.GsA);
    else if (E.ellA) std::snprintf(names[1], 48, "ell_dual_step<%s%s>", pt, ra);
    else if (E.stageA)
        std::snprintf(names[1], 48, "staged_dual_step<%s%s,%d>", pt, E.w16A ? ",w16" : ra, E.LA().threads);
    else if (E.streamA) std::snprintf(names[1], 48, "stream_dual_step<%s%s>", pt, ra);
    else std::snprintf(names[1], 48, "dual_step<%s,G=%d>", pt, E.GA);
    const double us[2] = {up, ud};
    const double bytes[2] = {b_primal, b_dual};
    int k = 0;
    for (int q = 0; q < 2 && k < cap; ++q, ++k) {
        std::snprintf(out[k].name, sizeof(out[k].name), "%s", names[q]);
        out[k].avg_us = us[q];
        out[k].launches = reps;
        out[k].bytes_per_launch = bytes[q];
    }
    *written = k;
    return VJ_OK;
}

}  // namespace vj
