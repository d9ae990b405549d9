// Kernel-level operators on caller device memory (vj_op_* in include/vajra_pdhg.h).
// These are the hot-loop kernels in isolation, so that each can be checked against the
// reference kernel contracts of src/solvers/gpu/spmv_kernels.cuh one by one.
#include <hip/hip_runtime.h>

#include <mutex>

#include "vj_device.hpp"
#include "vj_internal.hpp"

namespace vj {
namespace {

// Scratch of vj_op_diagnostics_static: the per-workgroup partials (4 doubles each) of at
// most kDiagMaxBlocks workgroups, so max(m, n) <= kDiagMaxBlocks * kBlock (8,388,608).
// Each GPU holds its own copy (1 MiB) until the process ends.
constexpr int kDiagMaxBlocks = 32768;
static __device__ __align__(256) double scratch[4 * kDiagMaxBlocks];

// Device address of `scratch` on each device, looked up once per device with that device
// current; host code passes this address to kernels, never the host-side symbol.
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_dev[kMaxDevices] = {};

int scratch_on_current_device(double** out) {
    int dev = 0;
    VJ_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) {
        set_error("device index " + std::to_string(dev) + " beyond the " + std::to_string(kMaxDevices) +
                  " devices the static scratch table covers");
        return VJ_ERR_UNSUPPORTED;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void* p = nullptr;
        VJ_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch_dev[dev] = static_cast<double*>(p);
    }
    *out = g_scratch_dev[dev];
    return VJ_OK;
}

// Row-group SpMV over int64 row pointers (the reference's CSR layout).
template <int G>
__global__ __launch_bounds__(kBlock) void op_spmv_groups(int rows, const int64_t* __restrict__ rp,
                                                         const int32_t* __restrict__ ci,
                                                         const double* __restrict__ av,
                                                         const double* __restrict__ x, double* __restrict__ y) {
    const int r = blockIdx.x * (kBlock / G) + threadIdx.x / G;
    const int lane = threadIdx.x % G;
    double acc = 0.0;
    if (r < rows)
        for (int64_t p = rp[r] + lane; p < rp[r + 1]; p += G) acc += av[p] * x[ci[p]];
    acc = group_sum<G>(acc);
    if (r < rows && lane == 0) y[r] = acc;
}

// One lane per row, ascending sum, no FMA; optional skip of |x_j| <= 1e-11.
__global__ __launch_bounds__(kBlock) void op_spmv_exact(int rows, const int64_t* __restrict__ rp,
                                                        const int32_t* __restrict__ ci,
                                                        const double* __restrict__ av,
                                                        const double* __restrict__ x, double* __restrict__ y,
                                                        int skip_tiny) {
#pragma clang fp contract(off)
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= rows) return;
    double acc = 0.0;
    for (int64_t p = rp[r]; p < rp[r + 1]; ++p) {
        const double xj = x[ci[p]];
        if (skip_tiny && fabs(xj) <= 1e-11) continue;
        const double prod = av[p] * xj;
        acc = acc + prod;
    }
    y[r] = acc;
}

__global__ void op_dual(int m, double sigma, const double* __restrict__ w, const double* __restrict__ lr,
                        const double* __restrict__ ur, double* __restrict__ y) {
#pragma clang fp contract(off)
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const double sw = sigma * w[i];
    const double y_hat = y[i] + sw;
    const double val = y_hat / sigma;
    const double clamped = (val < lr[i]) ? lr[i] : ((val > ur[i]) ? ur[i] : val);
    const double sc = sigma * clamped;
    y[i] = y_hat - sc;
}

__global__ void op_primal(int n, double tau, const double* __restrict__ v, const double* __restrict__ c,
                          const double* __restrict__ lc, const double* __restrict__ uc, double* __restrict__ x,
                          double* __restrict__ xb) {
#pragma clang fp contract(off)
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const double x_old = x[j];
    const double grad = v[j] + c[j];
    const double tg = tau * grad;
    double x_new = x_old - tg;
    if (x_new < lc[j]) x_new = lc[j];
    if (x_new > uc[j]) x_new = uc[j];
    x[j] = x_new;
    const double twice = 2.0 * x_new;
    xb[j] = twice - x_old;
}

// Block maxima / sums; final reduction in a second single-block launch (deterministic).
__global__ __launch_bounds__(kBlock) void op_diag_blocks(int m, int n, const double* __restrict__ ax,
                                                         const double* __restrict__ lr, const double* __restrict__ ur,
                                                         const double* __restrict__ aty, const double* __restrict__ c,
                                                         const double* __restrict__ x, const double* __restrict__ lc,
                                                         const double* __restrict__ uc, double* __restrict__ part) {
    __shared__ double s[4][kBlock];
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    double pv = 0.0, dv = 0.0, obj = 0.0, bad = 0.0;
    if (k < m) {
        const double a = ax[k];
        if (a < lr[k]) pv = lr[k] - a;
        if (a > ur[k] && a - ur[k] > pv) pv = a - ur[k];
        if (isnan(a) || isinf(a)) bad = 1.0;
    }
    if (k < n) {
        const double xj = x[k], lj = lc[k], uj = uc[k];
        if (xj < lj && lj - xj > pv) pv = lj - xj;
        if (xj > uj && xj - uj > pv) pv = xj - uj;
        const double dj = c[k] + aty[k];
        if (xj > lj + 1e-5 && dj > 1e-6 && dj > dv) dv = dj;
        if (xj < uj - 1e-5 && dj < -1e-6 && -dj > dv) dv = -dj;
        if (lj <= -1e20 && uj >= 1e20 && fabs(dj) > dv) dv = fabs(dj);
        obj = c[k] * xj;
        if (isnan(xj) || isinf(xj) || isnan(dj) || isinf(dj)) bad = 1.0;
    }
    s[0][threadIdx.x] = pv;
    s[1][threadIdx.x] = dv;
    s[2][threadIdx.x] = obj;
    s[3][threadIdx.x] = bad;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if (static_cast<int>(threadIdx.x) < w) {
            const int t = threadIdx.x;
            s[0][t] = s[0][t] < s[0][t + w] ? s[0][t + w] : s[0][t];
            s[1][t] = s[1][t] < s[1][t + w] ? s[1][t + w] : s[1][t];
            s[2][t] += s[2][t + w];
            s[3][t] = s[3][t] < s[3][t + w] ? s[3][t + w] : s[3][t];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int q = 0; q < 4; ++q) part[4 * blockIdx.x + q] = s[q][0];
}

__global__ __launch_bounds__(kBlock) void op_diag_final(int nb, const double* __restrict__ part, double* __restrict__ out) {
    __shared__ double s[4][kBlock];
    double a[4] = {0.0, 0.0, 0.0, 0.0};
    for (int b = threadIdx.x; b < nb; b += kBlock) {
        a[0] = a[0] < part[4 * b] ? part[4 * b] : a[0];
        a[1] = a[1] < part[4 * b + 1] ? part[4 * b + 1] : a[1];
        a[2] += part[4 * b + 2];
        a[3] = a[3] < part[4 * b + 3] ? part[4 * b + 3] : a[3];
    }
    for (int q = 0; q < 4; ++q) s[q][threadIdx.x] = a[q];
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if (static_cast<int>(threadIdx.x) < w) {
            const int t = threadIdx.x;
            s[0][t] = s[0][t] < s[0][t + w] ? s[0][t + w] : s[0][t];
            s[1][t] = s[1][t] < s[1][t + w] ? s[1][t + w] : s[1][t];
            s[2][t] += s[2][t + w];
            s[3][t] = s[3][t] < s[3][t + w] ? s[3][t + w] : s[3][t];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int q = 0; q < 4; ++q) out[q] = s[q][0];
}

}  // namespace
}  // namespace vj

using namespace vj;

extern "C" {

int vj_op_spmv(int rows, const int64_t* row_ptr, const int32_t* col_idx, const double* vals, const double* x,
               double* y, int lanes, int exact_order, int skip_tiny, void* stream) {
    clear_error();
    if (rows < 0 || !row_ptr || (rows > 0 && (!y || !x))) return VJ_ERR_INVALID_ARGUMENT;
    if (rows == 0) return VJ_OK;
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (exact_order) {
        hipLaunchKernelGGL(op_spmv_exact, dim3(grid_for(rows, kBlock)), dim3(kBlock), 0, s, rows, row_ptr, col_idx, vals,
                           x, y, skip_tiny);
    } else {
        switch (lanes) {
#define VJ_OP_CASE(G)                                                                                          \
    case G:                                                                                                    \
        hipLaunchKernelGGL((op_spmv_groups<G>), dim3(grid_for(rows, kBlock / G)), dim3(kBlock), 0, s, rows, row_ptr, \
                           col_idx, vals, x, y);                                                               \
        break;
            VJ_OP_CASE(2)
            VJ_OP_CASE(4)
            VJ_OP_CASE(8)
            VJ_OP_CASE(16)
            VJ_OP_CASE(32)
            VJ_OP_CASE(64)
#undef VJ_OP_CASE
            default:
                set_error("lanes must be one of 2, 4, 8, 16, 32, 64 (or use exact_order)");
                return VJ_ERR_INVALID_ARGUMENT;
        }
    }
    VJ_LAUNCHED();
    return VJ_OK;
}

int vj_op_dual_update(int m, double sigma, const double* w, const double* row_lower, const double* row_upper, double* y,
                      void* stream) {
    clear_error();
    if (m < 0 || !(sigma > 0.0)) return VJ_ERR_INVALID_ARGUMENT;
    if (m == 0) return VJ_OK;
    hipLaunchKernelGGL(op_dual, dim3(grid_for(m, kBlock)), dim3(kBlock), 0, static_cast<hipStream_t>(stream), m, sigma, w,
                       row_lower, row_upper, y);
    VJ_LAUNCHED();
    return VJ_OK;
}

int vj_op_primal_update(int n, double tau, const double* v, const double* c, const double* col_lower,
                        const double* col_upper, double* x, double* x_bar, void* stream) {
    clear_error();
    if (n < 0 || !(tau > 0.0)) return VJ_ERR_INVALID_ARGUMENT;
    if (n == 0) return VJ_OK;
    hipLaunchKernelGGL(op_primal, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, static_cast<hipStream_t>(stream), n, tau, v,
                       c, col_lower, col_upper, x, x_bar);
    VJ_LAUNCHED();
    return VJ_OK;
}

int vj_op_diagnostics(int m, int n, const double* Ax, const double* row_lower, const double* row_upper, const double* Aty,
                      const double* c, const double* x, const double* col_lower, const double* col_upper, double* out4,
                      void* stream) {
    clear_error();
    if (m < 0 || n < 0 || !out4) return VJ_ERR_INVALID_ARGUMENT;
    hipStream_t s = static_cast<hipStream_t>(stream);
    const int nb = grid_for(m > n ? m : n, kBlock);
    DevBuf part, o;
    VJ_TRY(part.alloc(static_cast<size_t>(nb) * 4 * 8));
    VJ_TRY(o.alloc(4 * 8));
    hipLaunchKernelGGL(op_diag_blocks, dim3(nb), dim3(kBlock), 0, s, m, n, Ax, row_lower, row_upper, Aty, c, x, col_lower,
                       col_upper, part.as<double>());
    hipLaunchKernelGGL(op_diag_final, dim3(1), dim3(kBlock), 0, s, nb, part.as<double>(), o.as<double>());
    VJ_LAUNCHED();
    VJ_HIP(hipMemcpyAsync(out4, o.p, 32, hipMemcpyDeviceToHost, s));
    VJ_HIP(hipStreamSynchronize(s));
    return VJ_OK;
}

int vj_op_diagnostics_static(int m, int n, const double* Ax, const double* row_lower, const double* row_upper,
                             const double* Aty, const double* c, const double* x, const double* col_lower,
                             const double* col_upper, double* out4_device, void* stream) {
    clear_error();
    if (m < 0 || n < 0 || !out4_device) return VJ_ERR_INVALID_ARGUMENT;
    const int nb = grid_for(m > n ? m : n, kBlock);
    if (nb > kDiagMaxBlocks) {
        set_error("vj_op_diagnostics_static: max(m, n) = " + std::to_string(m > n ? m : n) + " exceeds the " +
                  std::to_string(static_cast<int64_t>(kDiagMaxBlocks) * kBlock) +
                  " entries its static scratch covers; use vj_op_diagnostics");
        return VJ_ERR_INVALID_ARGUMENT;
    }
    double* part = nullptr;
    VJ_TRY(scratch_on_current_device(&part));
    hipStream_t s = static_cast<hipStream_t>(stream);
    // every partial the final reduction reads is written by this call's first kernel
    hipLaunchKernelGGL(op_diag_blocks, dim3(nb), dim3(kBlock), 0, s, m, n, Ax, row_lower, row_upper, Aty, c, x, col_lower,
                       col_upper, part);
    hipLaunchKernelGGL(op_diag_final, dim3(1), dim3(kBlock), 0, s, nb, part, out4_device);
    VJ_LAUNCHED();
    return VJ_OK;
}

}  // extern "C"
