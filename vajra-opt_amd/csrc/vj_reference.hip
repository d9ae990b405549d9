// Reference iteration map on the GPU (VJ_MODE_REFERENCE).
//
// Restates CpuPdhgSolver::solve (src/solvers/pdhg/pdhg.cpp:70-440) with the device doing
// the iterations: one lane per row, ascending-index sums and no fused multiply-add
// (#pragma clang fp contract(off)), so every SpMV, prox and extrapolation rounds exactly
// as the reference CPU code does. The per-check sums that the reference accumulates in
// index order (dual and primal objective, pdhg.cpp:280-319) are formed by one lane in
// that order from LDS-staged terms; maxima are order-free. The host reads a 48-byte
// diagnostics record per check and applies the reference's decisions (best iterate,
// stopping rule, step rebalancing, stagnation restart), mirroring the device seam of
// src/solvers/gpu/pdhg_cuda.cu:261-313. Setup runs on the host as in pdhg_cuda.cu:149-162:
// Ruiz (scaling.cpp:7-153), CSC->CSR (sparse.cpp:147-184), power iteration (pdhg.cpp:11-56).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "vj_context.hpp"
#include "vj_device.hpp"

namespace vj {

namespace {

using clk = std::chrono::steady_clock;
double secs(clk::time_point a, clk::time_point b) { return std::chrono::duration<double>(b - a).count(); }

constexpr double kZeroDrop = 1e-11;

// ---- device kernels -----------------------------------------------------------------
// w = A x_bar (skip |x_j| <= 1e-11, sparse.cpp:207), then the dual prox (pdhg.cpp:203-214).
__global__ __launch_bounds__(kBlock) void ref_dual(int m, const int* __restrict__ rp, const int* __restrict__ ci,
                                                   const double* __restrict__ av, const double* __restrict__ xb,
                                                   double sigma, const double* __restrict__ lr,
                                                   const double* __restrict__ ur, double* __restrict__ y,
                                                   int* __restrict__ nan_flag) {
#pragma clang fp contract(off)
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    double w = 0.0;
    for (int p = rp[i]; p < rp[i + 1]; ++p) {
        const double xj = xb[ci[p]];
        if (fabs(xj) <= kZeroDrop) continue;
        const double prod = (1.0 * xj) * av[p];
        w = w + prod;
    }
    const double sw = sigma * w;
    const double y_hat = y[i] + sw;
    const double li = lr[i], ui = ur[i];
    const double val = y_hat / sigma;
    const double clamped = (val < li) ? li : ((val > ui) ? ui : val);
    const double sc = sigma * clamped;
    const double y_new = y_hat - sc;
    if (isnan(y_new) || isinf(y_new)) *nan_flag = 1;
    y[i] = y_new;
}

// v = A^T y (ascending rows, sparse.cpp:219-238), primal prox + extrapolation (pdhg.cpp:228-241).
__global__ __launch_bounds__(kBlock) void ref_primal(int n, const int* __restrict__ tp, const int* __restrict__ ti,
                                                     const double* __restrict__ tv, const double* __restrict__ y,
                                                     double tau, const double* __restrict__ c,
                                                     const double* __restrict__ lc, const double* __restrict__ uc,
                                                     double* __restrict__ x, double* __restrict__ xb,
                                                     double* __restrict__ v_out, int* __restrict__ nan_flag) {
#pragma clang fp contract(off)
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    double dot = 0.0;
    for (int p = tp[j]; p < tp[j + 1]; ++p) {
        const double prod = tv[p] * y[ti[p]];
        dot = dot + prod;
    }
    const double v = 0.0 + 1.0 * dot;
    v_out[j] = v;
    const double x_old = x[j];
    const double grad = v + c[j];
    const double tg = tau * grad;
    double x_new = x_old - tg;
    if (x_new < lc[j]) x_new = lc[j];
    if (x_new > uc[j]) x_new = uc[j];
    if (isnan(x_new) || isinf(x_new)) *nan_flag = 1;
    x[j] = x_new;
    const double twice = 2.0 * x_new;
    xb[j] = twice - x_old;
}

// Ax for the check (pdhg.cpp:260), with the same skip rule.
__global__ __launch_bounds__(kBlock) void ref_spmv_skip(int m, const int* __restrict__ rp, const int* __restrict__ ci,
                                                        const double* __restrict__ av, const double* __restrict__ x,
                                                        double* __restrict__ ax) {
#pragma clang fp contract(off)
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    double w = 0.0;
    for (int p = rp[i]; p < rp[i + 1]; ++p) {
        const double xj = x[ci[p]];
        if (fabs(xj) <= kZeroDrop) continue;
        const double prod = (1.0 * xj) * av[p];
        w = w + prod;
    }
    ax[i] = w;
}

// Per-element check terms (pdhg.cpp:261-319). Non-contributing entries store 0, which is
// exact for the ordered sums (x - 0 = x; the running sums never become -0).
// Block maxima of the primal and dual violations go to pmax/dmax[blockIdx].
__global__ __launch_bounds__(kBlock) void ref_check_terms(int m, int n, const double* __restrict__ ax,
                                                          const double* __restrict__ lr, const double* __restrict__ ur,
                                                          const double* __restrict__ y, const double* __restrict__ x,
                                                          const double* __restrict__ lc, const double* __restrict__ uc,
                                                          const double* __restrict__ c, const double* __restrict__ v,
                                                          double* __restrict__ term_r, double* __restrict__ term_c,
                                                          double* __restrict__ term_p, double* __restrict__ pmax,
                                                          double* __restrict__ dmax) {
#pragma clang fp contract(off)
    __shared__ double lp[kBlock], ld[kBlock];
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    double pv = 0.0, dv = 0.0;
    if (k < m) {
        const double axi = ax[k], li = lr[k], ui = ur[k];
        if (axi < li) pv = li - axi;
        if (axi > ui) pv = (pv < axi - ui) ? axi - ui : pv;
        const double yi = y[k];
        double t = 0.0;
        if (yi > 1e-12 && ui < 1e20) t = yi * ui;
        else if (yi < -1e-12 && li > -1e20) t = yi * li;
        term_r[k] = t;
    }
    if (k < n) {
        const double xj = x[k], lj = lc[k], uj = uc[k];
        if (xj < lj) pv = (pv < lj - xj) ? lj - xj : pv;
        if (xj > uj) pv = (pv < xj - uj) ? xj - uj : pv;
        const double dj = c[k] + v[k];
        if (xj > lj + 1e-5 && dj > 1e-6) dv = (dv < dj) ? dj : dv;
        if (xj < uj - 1e-5 && dj < -1e-6) dv = (dv < -dj) ? -dj : dv;
        if (lj <= -1e20 && uj >= 1e20) dv = (dv < fabs(dj)) ? fabs(dj) : dv;
        double t = 0.0;
        if (dj > 1e-12 && lj > -1e20) t = dj * lj;
        else if (dj < -1e-12 && uj < 1e20) t = dj * uj;
        term_c[k] = t;
        term_p[k] = c[k] * xj;
    }
    lp[threadIdx.x] = pv;
    ld[threadIdx.x] = dv;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if (static_cast<int>(threadIdx.x) < w) {
            const double a = lp[threadIdx.x + w], b = ld[threadIdx.x + w];
            if (lp[threadIdx.x] < a) lp[threadIdx.x] = a;
            if (ld[threadIdx.x] < b) ld[threadIdx.x] = b;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        pmax[blockIdx.x] = lp[0];
        dmax[blockIdx.x] = ld[0];
    }
}

struct RefDiag {   // 48 bytes, read by the host once per check
    double primal_residual;
    double dual_residual;
    double primal_obj;
    double dual_obj;
    double gap;
    int32_t nan_flag;
    int32_t pad;
};

// One workgroup: maxima over blocks, then the ordered sums in one lane over LDS chunks.
__global__ __launch_bounds__(kBlock) void ref_check_final(int m, int n, int nblocks, const double* __restrict__ pmax,
                                                          const double* __restrict__ dmax,
                                                          const double* __restrict__ term_r,
                                                          const double* __restrict__ term_c,
                                                          const double* __restrict__ term_p,
                                                          const int* __restrict__ nan_flag, RefDiag* __restrict__ out) {
#pragma clang fp contract(off)
    constexpr int kChunk = 4096;
    __shared__ double buf[kChunk];
    __shared__ double red[2][kBlock];
    double pv = 0.0, dv = 0.0;
    for (int b = threadIdx.x; b < nblocks; b += kBlock) {
        if (pv < pmax[b]) pv = pmax[b];
        if (dv < dmax[b]) dv = dmax[b];
    }
    red[0][threadIdx.x] = pv;
    red[1][threadIdx.x] = dv;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if (static_cast<int>(threadIdx.x) < w) {
            if (red[0][threadIdx.x] < red[0][threadIdx.x + w]) red[0][threadIdx.x] = red[0][threadIdx.x + w];
            if (red[1][threadIdx.x] < red[1][threadIdx.x + w]) red[1][threadIdx.x] = red[1][threadIdx.x + w];
        }
        __syncthreads();
    }
    double dual_obj = 0.0, primal_obj = 0.0;
    // rows: dual_obj -= t_i (ascending i)
    for (int base = 0; base < m; base += kChunk) {
        const int len = (m - base) < kChunk ? (m - base) : kChunk;
        for (int q = threadIdx.x; q < len; q += kBlock) buf[q] = term_r[base + q];
        __syncthreads();
        if (threadIdx.x == 0)
            for (int q = 0; q < len; ++q) dual_obj = dual_obj - buf[q];
        __syncthreads();
    }
    // columns: dual_obj += t_j (ascending j)
    for (int base = 0; base < n; base += kChunk) {
        const int len = (n - base) < kChunk ? (n - base) : kChunk;
        for (int q = threadIdx.x; q < len; q += kBlock) buf[q] = term_c[base + q];
        __syncthreads();
        if (threadIdx.x == 0)
            for (int q = 0; q < len; ++q) dual_obj = dual_obj + buf[q];
        __syncthreads();
    }
    for (int base = 0; base < n; base += kChunk) {
        const int len = (n - base) < kChunk ? (n - base) : kChunk;
        for (int q = threadIdx.x; q < len; q += kBlock) buf[q] = term_p[base + q];
        __syncthreads();
        if (threadIdx.x == 0)
            for (int q = 0; q < len; ++q) primal_obj = primal_obj + buf[q];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        RefDiag d;
        d.primal_residual = red[0][0];
        d.dual_residual = red[1][0];
        d.primal_obj = primal_obj;
        d.dual_obj = dual_obj;
        d.gap = fabs(primal_obj - dual_obj) / (1.0 + fabs(primal_obj) + fabs(dual_obj));
        d.nan_flag = *nan_flag;
        d.pad = 0;
        *out = d;
    }
}

// ---- host setup (restatements; compiled without FMA contraction on x86-64) ------------
double host_max(double a, double b) { return (a < b) ? b : a; }

// scaling.cpp:7-153 on CSC values; returns cumulative row/col scales.
void host_ruiz(int m, int n, const std::vector<int64_t>& cp, const std::vector<int32_t>& ri,
               std::vector<double>& vals, std::vector<double>& c, std::vector<double>& rl,
               std::vector<double>& ru, std::vector<double>& cl, std::vector<double>& cu,
               std::vector<double>& rs, std::vector<double>& cs) {
    rs.assign(m, 1.0);
    cs.assign(n, 1.0);
    std::vector<double> rk(m, 1.0), ck(n, 1.0), rinf(m), cinf(n);
    for (int iter = 0; iter < 10; ++iter) {
        std::fill(rinf.begin(), rinf.end(), 0.0);
        for (int j = 0; j < n; ++j)
            for (int64_t k = cp[j]; k < cp[j + 1]; ++k) rinf[ri[k]] = host_max(rinf[ri[k]], std::fabs(vals[k]));
        for (int j = 0; j < n; ++j) {
            cinf[j] = 0.0;
            for (int64_t k = cp[j]; k < cp[j + 1]; ++k) cinf[j] = host_max(cinf[j], std::fabs(vals[k]));
        }
        double er = 0.0, ec = 0.0;
        for (int i = 0; i < m; ++i)
            if (rinf[i] > kZeroDrop) er = host_max(er, std::fabs(1.0 - rinf[i]));
        for (int j = 0; j < n; ++j)
            if (cinf[j] > kZeroDrop) ec = host_max(ec, std::fabs(1.0 - cinf[j]));
        if (er <= 0.05 && ec <= 0.05) break;
        for (int i = 0; i < m; ++i) {
            rk[i] = rinf[i] > kZeroDrop ? 1.0 / std::sqrt(rinf[i]) : 1.0;
            rs[i] *= rk[i];
        }
        for (int j = 0; j < n; ++j) {
            ck[j] = cinf[j] > kZeroDrop ? 1.0 / std::sqrt(cinf[j]) : 1.0;
            cs[j] *= ck[j];
        }
        for (int j = 0; j < n; ++j) {
            const double cj = ck[j];
            for (int64_t k = cp[j]; k < cp[j + 1]; ++k) {
                const double f = rk[ri[k]] * cj;
                vals[k] *= f;
            }
        }
    }
    for (int j = 0; j < n; ++j) c[j] *= cs[j];
    for (int i = 0; i < m; ++i) if (rl[i] > -1e20) rl[i] *= rs[i];
    for (int i = 0; i < m; ++i) if (ru[i] < 1e20) ru[i] *= rs[i];
    for (int j = 0; j < n; ++j) {
        const double inv = 1.0 / cs[j];
        if (cl[j] > -1e20) cl[j] *= inv;
    }
    for (int j = 0; j < n; ++j) {
        const double inv = 1.0 / cs[j];
        if (cu[j] < 1e20) cu[j] *= inv;
    }
}

// sparse.cpp:196-238 with alpha = 1, beta = 0
void host_mult(int m, int n, const std::vector<int64_t>& cp, const std::vector<int32_t>& ri,
               const std::vector<double>& v, const std::vector<double>& x, std::vector<double>& y) {
    std::fill(y.begin(), y.end(), 0.0);
    for (int j = 0; j < n; ++j) {
        const double xj = x[j];
        if (std::fabs(xj) <= kZeroDrop) continue;
        const double a = 1.0 * xj;
        for (int64_t k = cp[j]; k < cp[j + 1]; ++k) {
            const double p = a * v[k];
            y[ri[k]] = y[ri[k]] + p;
        }
    }
    (void)m;
}
void host_mult_t(int n, const std::vector<int64_t>& cp, const std::vector<int32_t>& ri,
                 const std::vector<double>& v, const std::vector<double>& y, std::vector<double>& out) {
    for (int j = 0; j < n; ++j) {
        double dot = 0.0;
        for (int64_t k = cp[j]; k < cp[j + 1]; ++k) {
            const double p = v[k] * y[ri[k]];
            dot = dot + p;
        }
        const double s = 1.0 * dot;
        out[j] = 0.0 + s;
    }
}

// pdhg.cpp:11-56
double host_spectral_norm(int m, int n, const std::vector<int64_t>& cp, const std::vector<int32_t>& ri,
                          const std::vector<double>& v) {
    if (m == 0 || n == 0 || cp[n] == 0) return 1.0;
    std::vector<double> vv(n, 1.0 / std::sqrt(static_cast<double>(n))), w(m), u(n);
    double lambda_prev = 0.0, lambda = 1.0;
    for (int iter = 0; iter < 25; ++iter) {
        host_mult(m, n, cp, ri, v, vv, w);
        host_mult_t(n, cp, ri, v, w, u);
        double ns = 0.0;
        for (int j = 0; j < n; ++j) {
            const double p = u[j] * u[j];
            ns = ns + p;
        }
        const double norm = std::sqrt(ns);
        if (norm < 1e-15 || std::isnan(norm) || std::isinf(norm)) return (lambda > 1e-6) ? std::sqrt(lambda) : 1e-3;
        lambda = norm;
        if (iter > 0 && std::fabs(lambda - lambda_prev) <= 1e-4 * lambda) break;
        lambda_prev = lambda;
        const double inv = 1.0 / norm;
        for (int j = 0; j < n; ++j) vv[j] = u[j] * inv;
    }
    const double smax = std::sqrt(lambda);
    return (smax > 1e-12) ? smax : 1e-6;
}

}  // namespace

struct RefEngine {
    DevBuf rp, ci, av, tp, ti, tv, c, lc, uc, lr, ur;
    DevBuf x, xb, y, v, ax, best_x, best_y, term_r, term_c, term_p, pmax, dmax, nan_flag, diag;
    RefDiag* h_diag = nullptr;
    ~RefEngine() {
        if (h_diag) (void)hipHostFree(h_diag);
    }
};
void destroy_ref(RefEngine* e) { delete e; }

int reference_solve(vj_ctx* ctx, vj_result* out, int64_t stop_at_iter, double* traj_x, double* traj_xbar,
                    double* traj_y) {
    const auto start = clk::now();
    Problem& P = ctx->prob;
    const vj_params& prm = ctx->params;
    if (P.count != 1) {
        set_error("the reference iteration map solves a single LP");
        return VJ_ERR_UNSUPPORTED;
    }
    const int m = P.m, n = P.n;
    const bool is_max = P.maximize_i[0] != 0;
    hipStream_t s = ctx->stream;
    std::memset(out, 0, sizeof(*out));

    // setup on the host (pdhg.cpp:129-184)
    std::vector<double> vals = P.h_tv;   // CSC values
    std::vector<double> wc = P.h_c, wrl = P.h_lr, wru = P.h_ur, wcl = P.h_lc, wcu = P.h_uc;
    if (is_max)
        for (double& cj : wc) cj = -cj;
    std::vector<double> rs, cs;
    const auto setup0 = clk::now();
    bool scaled = false;
    if (prm.enable_scaling) {
        host_ruiz(m, n, P.h_tp, P.h_ti, vals, wc, wrl, wru, wcl, wcu, rs, cs);
        scaled = true;
    }
    // to_csr of the scaled matrix (sparse.cpp:147-184)
    std::vector<int32_t> rp(m + 1, 0), ci(P.nnz);
    std::vector<double> av(P.nnz);
    {
        std::vector<int64_t> cnt(m + 1, 0);
        for (int64_t k = 0; k < P.nnz; ++k) cnt[P.h_ti[k] + 1]++;
        for (int i = 0; i < m; ++i) cnt[i + 1] += cnt[i];
        for (int i = 0; i <= m; ++i) rp[i] = static_cast<int32_t>(cnt[i]);
        for (int j = 0; j < n; ++j)
            for (int64_t k = P.h_tp[j]; k < P.h_tp[j + 1]; ++k) {
                const int64_t d = cnt[P.h_ti[k]]++;
                ci[d] = j;
                av[d] = vals[k];
            }
    }
    const double L = host_spectral_norm(m, n, P.h_tp, P.h_ti, vals);
    double safety = prm.step_size_safety;
    safety = (safety < 0.5) ? 0.5 : ((0.99 < safety) ? 0.99 : safety);
    double tau = safety / L, sigma = safety / L;
    std::vector<double> x0(n);
    for (int j = 0; j < n; ++j) {
        double val = 0.0;
        if (val < wcl[j]) val = wcl[j];
        if (val > wcu[j]) val = wcu[j];
        x0[j] = val;
    }
    const double setup_time = secs(setup0, clk::now());

    // upload (pdhg_cuda.cu:187-231 counterpart)
    const auto up0 = clk::now();
    if (!ctx->ref) ctx->ref.reset(new RefEngine());
    RefEngine& R = *ctx->ref;
    int64_t transfers = 0;
    auto up = [&](DevBuf& b, const void* h, size_t bytes) -> int {
        VJ_TRY(b.alloc(bytes));
        VJ_HIP(hipMemcpyAsync(b.p, h, bytes, hipMemcpyHostToDevice, s));
        ++transfers;
        return VJ_OK;
    };
    std::vector<int32_t> tp32(n + 1);
    for (int j = 0; j <= n; ++j) tp32[j] = static_cast<int32_t>(P.h_tp[j]);
    VJ_TRY(up(R.rp, rp.data(), (m + 1) * 4));
    VJ_TRY(up(R.ci, ci.data(), P.nnz * 4));
    VJ_TRY(up(R.av, av.data(), P.nnz * 8));
    VJ_TRY(up(R.tp, tp32.data(), (n + 1) * 4));
    VJ_TRY(up(R.ti, P.h_ti.data(), P.nnz * 4));
    VJ_TRY(up(R.tv, vals.data(), P.nnz * 8));
    VJ_TRY(up(R.c, wc.data(), n * 8));
    VJ_TRY(up(R.lc, wcl.data(), n * 8));
    VJ_TRY(up(R.uc, wcu.data(), n * 8));
    VJ_TRY(up(R.lr, wrl.data(), m * 8));
    VJ_TRY(up(R.ur, wru.data(), m * 8));
    VJ_TRY(up(R.x, x0.data(), n * 8));
    VJ_TRY(up(R.xb, x0.data(), n * 8));
    VJ_TRY(up(R.best_x, x0.data(), n * 8));
    std::vector<double> zeros_m(m, 0.0);
    VJ_TRY(up(R.y, zeros_m.data(), m * 8));
    VJ_TRY(up(R.best_y, zeros_m.data(), m * 8));
    const int gmax = grid_for(std::max(m, n), kBlock);
    for (DevBuf* b : {&R.v, &R.term_c, &R.term_p}) VJ_TRY(b->alloc(static_cast<size_t>(n) * 8));
    for (DevBuf* b : {&R.ax, &R.term_r}) VJ_TRY(b->alloc(static_cast<size_t>(m) * 8));
    VJ_TRY(R.pmax.alloc(static_cast<size_t>(gmax) * 8));
    VJ_TRY(R.dmax.alloc(static_cast<size_t>(gmax) * 8));
    VJ_TRY(R.nan_flag.alloc(16));
    VJ_TRY(R.diag.alloc(sizeof(RefDiag)));
    int zero = 0;
    VJ_HIP(hipMemcpyAsync(R.nan_flag.p, &zero, 4, hipMemcpyHostToDevice, s));
    if (!R.h_diag) VJ_HIP(hipHostMalloc(reinterpret_cast<void**>(&R.h_diag), sizeof(RefDiag), 0));
    VJ_HIP(hipStreamSynchronize(s));
    const double upload_time = secs(up0, clk::now());

    const auto it0 = clk::now();
    const int64_t max_iter = prm.max_iterations;
    const int freq = std::max(1, prm.check_frequency);
    double best_inf = 1e20;
    int stag = 0, restarts = 0;
    bool converged = false;
    int status = VJ_STATUS_NOT_SOLVED;
    double last_p = 0, last_d = 0, last_g = 0;
    int64_t iter = 0;
    int64_t checks = 0;
    const int gm = grid_for(m, kBlock), gn = grid_for(n, kBlock);
    for (iter = 1; iter <= max_iter; ++iter) {
        hipLaunchKernelGGL(ref_dual, dim3(gm), dim3(kBlock), 0, s, m, R.rp.as<int>(), R.ci.as<int>(), R.av.as<double>(),
                           R.xb.as<double>(), sigma, R.lr.as<double>(), R.ur.as<double>(), R.y.as<double>(),
                           R.nan_flag.as<int>());
        hipLaunchKernelGGL(ref_primal, dim3(gn), dim3(kBlock), 0, s, n, R.tp.as<int>(), R.ti.as<int>(), R.tv.as<double>(),
                           R.y.as<double>(), tau, R.c.as<double>(), R.lc.as<double>(), R.uc.as<double>(),
                           R.x.as<double>(), R.xb.as<double>(), R.v.as<double>(), R.nan_flag.as<int>());
        VJ_LAUNCHED();
        if (stop_at_iter > 0 && iter == stop_at_iter) break;
        if (iter % freq == 0 || iter == max_iter) {
            if (secs(start, clk::now()) > prm.time_limit) {
                status = VJ_STATUS_TIME_LIMIT;
                break;
            }
            hipLaunchKernelGGL(ref_spmv_skip, dim3(gm), dim3(kBlock), 0, s, m, R.rp.as<int>(), R.ci.as<int>(),
                               R.av.as<double>(), R.x.as<double>(), R.ax.as<double>());
            hipLaunchKernelGGL(ref_check_terms, dim3(gmax), dim3(kBlock), 0, s, m, n, R.ax.as<double>(), R.lr.as<double>(),
                               R.ur.as<double>(), R.y.as<double>(), R.x.as<double>(), R.lc.as<double>(), R.uc.as<double>(),
                               R.c.as<double>(), R.v.as<double>(), R.term_r.as<double>(), R.term_c.as<double>(),
                               R.term_p.as<double>(), R.pmax.as<double>(), R.dmax.as<double>());
            hipLaunchKernelGGL(ref_check_final, dim3(1), dim3(kBlock), 0, s, m, n, gmax, R.pmax.as<double>(),
                               R.dmax.as<double>(), R.term_r.as<double>(), R.term_c.as<double>(), R.term_p.as<double>(),
                               R.nan_flag.as<int>(), R.diag.as<RefDiag>());
            VJ_LAUNCHED();
            VJ_HIP(hipMemcpyAsync(R.h_diag, R.diag.p, sizeof(RefDiag), hipMemcpyDeviceToHost, s));
            VJ_HIP(hipStreamSynchronize(s));
            ++transfers;
            ++checks;
            const RefDiag d = *R.h_diag;
            if (d.nan_flag) {
                status = VJ_STATUS_NUMERICAL_ERROR;
                break;
            }
            const double pr = d.primal_residual, dr = d.dual_residual, gap = d.gap;
            last_p = pr;
            last_d = dr;
            last_g = gap;
            const double total = pr + dr;
            if (total < best_inf) {
                best_inf = total;
                VJ_HIP(hipMemcpyAsync(R.best_x.p, R.x.p, n * 8, hipMemcpyDeviceToDevice, s));
                VJ_HIP(hipMemcpyAsync(R.best_y.p, R.y.p, m * 8, hipMemcpyDeviceToDevice, s));
                stag = 0;
            } else {
                stag += freq;
            }
            if (pr <= prm.tolerance && dr <= prm.tolerance && gap <= prm.gap_tolerance) {
                converged = true;
                VJ_HIP(hipMemcpyAsync(R.best_x.p, R.x.p, n * 8, hipMemcpyDeviceToDevice, s));
                VJ_HIP(hipMemcpyAsync(R.best_y.p, R.y.p, m * 8, hipMemcpyDeviceToDevice, s));
                status = VJ_STATUS_OPTIMAL;
                break;
            }
            if (pr > 10.0 * dr && dr > 0.0) {
                sigma = std::min(sigma * 1.3, 1e6);
                tau = (safety * safety) / (L * L * sigma);
            } else if (dr > 10.0 * pr && pr > 0.0) {
                tau = std::min(tau * 1.3, 1e6);
                sigma = (safety * safety) / (L * L * tau);
            }
            if (stag >= prm.restart_stagnation_limit) {
                VJ_HIP(hipMemcpyAsync(R.xb.p, R.x.p, n * 8, hipMemcpyDeviceToDevice, s));
                ++restarts;
                stag = 0;
            }
        }
    }
    // a NaN raised between checks is reported at the end (the reference stops at once)
    int nanf = 0;
    VJ_HIP(hipMemcpyAsync(&nanf, R.nan_flag.p, 4, hipMemcpyDeviceToHost, s));
    VJ_HIP(hipStreamSynchronize(s));
    if (nanf && status == VJ_STATUS_NOT_SOLVED) status = VJ_STATUS_NUMERICAL_ERROR;
    const double iter_time = secs(it0, clk::now());

    if (stop_at_iter > 0) {
        if (traj_x) VJ_HIP(hipMemcpy(traj_x, R.x.p, n * 8, hipMemcpyDeviceToHost));
        if (traj_xbar) VJ_HIP(hipMemcpy(traj_xbar, R.xb.p, n * 8, hipMemcpyDeviceToHost));
        if (traj_y) VJ_HIP(hipMemcpy(traj_y, R.y.p, m * 8, hipMemcpyDeviceToHost));
    }
    if (!converged && status == VJ_STATUS_NOT_SOLVED && iter >= max_iter)
        status = (best_inf <= 1e-4) ? VJ_STATUS_FEASIBLE : VJ_STATUS_ITERATION_LIMIT;

    // download the best iterate and unscale (pdhg.cpp:380-411)
    const auto dn0 = clk::now();
    std::vector<double> bx(n), by(m);
    VJ_HIP(hipMemcpy(bx.data(), R.best_x.p, n * 8, hipMemcpyDeviceToHost));
    VJ_HIP(hipMemcpy(by.data(), R.best_y.p, m * 8, hipMemcpyDeviceToHost));
    transfers += 2;
    if (scaled) {
        for (int j = 0; j < n; ++j) bx[j] *= cs[j];
        for (int i = 0; i < m; ++i) by[i] *= rs[i];
    }
    std::vector<double> ycan(m);
    ctx->ref_row_dual.assign(m, 0.0);
    for (int i = 0; i < m; ++i) {
        ycan[i] = -by[i];
        ctx->ref_row_dual[i] = is_max ? by[i] : -by[i];
    }
    VJ_TRY(ctx->x_sol.alloc(static_cast<size_t>(n) * 8));
    VJ_TRY(ctx->y_sol.alloc(static_cast<size_t>(m) * 8));
    VJ_HIP(hipMemcpy(ctx->x_sol.p, bx.data(), n * 8, hipMemcpyHostToDevice));
    VJ_HIP(hipMemcpy(ctx->y_sol.p, ycan.data(), m * 8, hipMemcpyHostToDevice));
    ctx->have_solution = true;
    const double download_time = secs(dn0, clk::now());
    double audit[8] = {0};
    VJ_TRY(audit_device(ctx, ctx->x_sol.as<double>(), ctx->y_sol.as<double>(), audit));

    out->status = status;
    out->converged = converged ? 1 : 0;
    out->iterations = iter;
    out->restart_count = restarts;
    out->primal_residual = last_p;
    out->dual_residual = last_d;
    out->duality_gap = last_g;
    out->rel_kkt = audit[0];
    out->rel_primal = audit[1];
    out->rel_dual = audit[2];
    out->rel_gap = audit[3];
    out->primal_objective = audit[4];
    out->dual_objective = audit[5];
    out->spectral_norm = L;
    out->tau = tau;
    out->sigma = sigma;
    out->primal_weight = 1.0;
    out->setup_time = setup_time;
    out->upload_time = upload_time;
    out->iteration_time = iter_time;
    out->download_time = download_time;
    out->total_time = secs(start, clk::now());
    out->time_to_target = converged ? iter_time : -1.0;
    out->host_device_transfers = transfers;
    out->lanes_A = 1;
    out->lanes_At = 1;
    out->check_interval = freq;
    out->checks = checks;
    return VJ_OK;
}

}  // namespace vj
