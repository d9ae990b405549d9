// Random-gather ceiling of one MI355X: how many independent 4- or 8-byte loads at random
// positions of a vector the chip serves per second, as a function of the vector's size
// (L2-resident, Infinity-Cache-resident, HBM), of the column split (workgroup w gathers
// only from block w mod S, the engine's XCD split) and of the loads each lane keeps in
// flight (U). This is the bound of the PDHG steps, whose every nonzero is one such gather
// (x_bar[j] in A x_bar, y[i] in A^T y); the HBM byte roofline is far above it.
//
//   gather_ceiling [reps]        -> CSV: type,vec_entries,vec_MB,S,U,gathers,avg_us,Ggathers_s,GB_s_useful
//   gather_ceiling [reps] spmv   -> CSV: mode,type,vec_entries,S,U,gathers,avg_us,Ggathers_s
//                                   (gathers alone / with a value stream / with 8-byte records)
//
// Every index is drawn on the host in [block_lo, block_hi) of its workgroup's block and
// checked against the vector length before the upload; the grid is gathers / (256 U)
// workgroups, the index buffer holds exactly gathers entries and the output buffer
// gathers entries (one per thread, at most one thread per gather).
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <algorithm>
#include <cstdlib>
#include <string>
#include <vector>

#define CK(x)                                                                              \
    do {                                                                                   \
        hipError_t e_ = (x);                                                               \
        if (e_ != hipSuccess) {                                                            \
            std::fprintf(stderr, "%s:%d %s: %s\n", __FILE__, __LINE__, #x, hipGetErrorString(e_)); \
            std::exit(2);                                                                  \
        }                                                                                  \
    } while (0)

constexpr int kThreads = 256;

template <class T, int U>
__global__ __launch_bounds__(kThreads) void gather_kernel(const int* __restrict__ idx, const T* __restrict__ vec,
                                                          T* __restrict__ out) {
    const int64_t base = static_cast<int64_t>(blockIdx.x) * kThreads * U + threadIdx.x;
    int c[U];
#pragma unroll
    for (int u = 0; u < U; ++u) c[u] = idx[base + u * kThreads];
    T v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) v[u] = vec[c[u]];
    T acc = 0;
#pragma unroll
    for (int u = 0; u < U; ++u) acc += v[u];
    out[static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x] = acc;
}

// The same gathers with the matrix value stream of an SpMV beside the index stream: either a
// separate value array (two 4-byte loads per entry) or interleaved (index, value) records
// of 8 bytes (one load per entry, FP32 values only).
template <class T, int U>
__global__ __launch_bounds__(kThreads) void gather_val_kernel(const int* __restrict__ idx, const T* __restrict__ val,
                                                              const T* __restrict__ vec, T* __restrict__ out) {
    const int64_t base = static_cast<int64_t>(blockIdx.x) * kThreads * U + threadIdx.x;
    int c[U];
    T a[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
        c[u] = __builtin_nontemporal_load(idx + base + u * kThreads);
        a[u] = __builtin_nontemporal_load(val + base + u * kThreads);
    }
    T acc = 0;
#pragma unroll
    for (int u = 0; u < U; ++u) acc += a[u] * vec[c[u]];
    out[static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x] = acc;
}

template <int U>
__global__ __launch_bounds__(kThreads) void gather_rec_kernel(const int2* __restrict__ rec, const float* __restrict__ vec,
                                                              float* __restrict__ out) {
    const int64_t base = static_cast<int64_t>(blockIdx.x) * kThreads * U + threadIdx.x;
    int2 r[U];
#pragma unroll
    for (int u = 0; u < U; ++u) r[u] = rec[base + u * kThreads];
    float acc = 0;
#pragma unroll
    for (int u = 0; u < U; ++u) acc += __int_as_float(r[u].y) * vec[r[u].x];
    out[static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x] = acc;
}

static uint64_t g_state = 0x5EED2026ULL;
static inline uint64_t next_u64() {  // SplitMix64
    uint64_t z = (g_state += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

template <class T, int U>
static double time_kernel(int64_t gathers, const int* d_idx, const T* d_vec, T* d_out, int reps, hipStream_t s) {
    const int grid = static_cast<int>(gathers / (kThreads * U));
    hipEvent_t a, b;
    CK(hipEventCreate(&a));
    CK(hipEventCreate(&b));
    for (int w = 0; w < 3; ++w) gather_kernel<T, U><<<grid, kThreads, 0, s>>>(d_idx, d_vec, d_out);
    CK(hipGetLastError());
    CK(hipEventRecord(a, s));
    for (int r = 0; r < reps; ++r) gather_kernel<T, U><<<grid, kThreads, 0, s>>>(d_idx, d_vec, d_out);
    CK(hipEventRecord(b, s));
    CK(hipEventSynchronize(b));
    float ms = 0.f;
    CK(hipEventElapsedTime(&ms, a, b));
    CK(hipEventDestroy(a));
    CK(hipEventDestroy(b));
    return 1e3 * ms / reps;
}

template <class F>
static double time_launch(F launch, int reps, hipStream_t s) {
    hipEvent_t a, b;
    CK(hipEventCreate(&a));
    CK(hipEventCreate(&b));
    for (int w = 0; w < 3; ++w) launch();
    CK(hipGetLastError());
    CK(hipEventRecord(a, s));
    for (int r = 0; r < reps; ++r) launch();
    CK(hipEventRecord(b, s));
    CK(hipEventSynchronize(b));
    float ms = 0.f;
    CK(hipEventElapsedTime(&ms, a, b));
    CK(hipEventDestroy(a));
    CK(hipEventDestroy(b));
    return 1e3 * ms / reps;
}

// "spmv" mode: gathers with the matrix streams of an SpMV, U = 4, on the vector sizes and
// splits of the 1M x 2M LP's two products (x_bar: 2M entries, y: 1M entries).
static int run_spmv_mode(int reps) {
    constexpr int U = 4;
    const int64_t gathers = 12582912;
    const int grid = static_cast<int>(gathers / (kThreads * U));
    hipStream_t s;
    CK(hipStreamCreate(&s));
    const int64_t nmax = 2097152;
    double *d_vec = nullptr, *d_val = nullptr, *d_out = nullptr;
    int* d_idx = nullptr;
    int2* d_rec = nullptr;
    CK(hipMalloc(&d_vec, nmax * sizeof(double)));
    CK(hipMalloc(&d_val, gathers * sizeof(double)));
    CK(hipMalloc(&d_out, gathers * sizeof(double)));
    CK(hipMalloc(&d_idx, gathers * sizeof(int)));
    CK(hipMalloc(&d_rec, gathers * sizeof(int2)));
    {
        std::vector<double> ones(std::max<int64_t>(nmax, gathers), 1.0);
        CK(hipMemcpy(d_vec, ones.data(), nmax * sizeof(double), hipMemcpyHostToDevice));
        CK(hipMemcpy(d_val, ones.data(), gathers * sizeof(double), hipMemcpyHostToDevice));
    }
    std::vector<int> idx(gathers);
    std::vector<int2> rec(gathers);
    std::printf("mode,type,vec_entries,S,U,gathers,avg_us,Ggathers_s\n");
    const int64_t ns[] = {2097152, 1048576};
    for (int64_t n : ns) {
        for (int S : {1, 2, 4, 8}) {
            const int64_t blk = n / S;
            for (int64_t q = 0; q < gathers; ++q) {
                const int64_t w = (q / kThreads) / U;
                idx[q] = static_cast<int>((w % S) * blk + static_cast<int64_t>(next_u64() % static_cast<uint64_t>(blk)));
                if (idx[q] < 0 || idx[q] >= n) {
                    std::fprintf(stderr, "index out of range\n");
                    return 3;
                }
                rec[q].x = idx[q];
                rec[q].y = __builtin_bit_cast(int, 1.0f);
            }
            CK(hipMemcpy(d_idx, idx.data(), gathers * sizeof(int), hipMemcpyHostToDevice));
            CK(hipMemcpy(d_rec, rec.data(), gathers * sizeof(int2), hipMemcpyHostToDevice));
            auto report = [&](const char* mode, const char* type, double us) {
                std::printf("%s,%s,%lld,%d,%d,%lld,%.2f,%.1f\n", mode, type, static_cast<long long>(n), S, U,
                            static_cast<long long>(gathers), us, gathers / (us * 1e-6) / 1e9);
                std::fflush(stdout);
            };
            float* vf = reinterpret_cast<float*>(d_vec);
            float* af = reinterpret_cast<float*>(d_val);
            float* of = reinterpret_cast<float*>(d_out);
            report("gather", "f32", time_launch([&] { gather_kernel<float, U><<<grid, kThreads, 0, s>>>(d_idx, vf, of); }, reps, s));
            report("gather+val", "f32", time_launch([&] { gather_val_kernel<float, U><<<grid, kThreads, 0, s>>>(d_idx, af, vf, of); }, reps, s));
            report("record", "f32", time_launch([&] { gather_rec_kernel<U><<<grid, kThreads, 0, s>>>(d_rec, vf, of); }, reps, s));
            report("gather", "f64", time_launch([&] { gather_kernel<double, U><<<grid, kThreads, 0, s>>>(d_idx, d_vec, d_out); }, reps, s));
            report("gather+val", "f64", time_launch([&] { gather_val_kernel<double, U><<<grid, kThreads, 0, s>>>(d_idx, d_val, d_vec, d_out); }, reps, s));
        }
    }
    CK(hipFree(d_vec));
    CK(hipFree(d_val));
    CK(hipFree(d_out));
    CK(hipFree(d_idx));
    CK(hipFree(d_rec));
    CK(hipStreamDestroy(s));
    return 0;
}

int main(int argc, char** argv) {
    const int reps = argc > 1 ? std::atoi(argv[1]) : 10;
    if (argc > 2 && std::string(argv[2]) == "spmv") return run_spmv_mode(reps);
    const int64_t gathers = 12582912;  // 3 * 2^22: a multiple of 256 * 16
    const int64_t sizes[] = {262144, 524288, 1048576, 2097152, 16777216, 134217728};
    const int splits[] = {1, 4, 8};
    hipStream_t s;
    CK(hipStreamCreate(&s));
    const int64_t nmax = sizes[sizeof(sizes) / sizeof(sizes[0]) - 1];
    double* d_vec = nullptr;
    double* d_out = nullptr;
    int* d_idx = nullptr;
    CK(hipMalloc(&d_vec, nmax * sizeof(double)));
    CK(hipMalloc(&d_out, gathers * sizeof(double)));  // U = 1 writes one entry per gather
    CK(hipMalloc(&d_idx, gathers * sizeof(int)));
    {
        std::vector<double> ones(nmax, 1.0);
        CK(hipMemcpy(d_vec, ones.data(), nmax * sizeof(double), hipMemcpyHostToDevice));
    }
    std::vector<int> idx(gathers);
    std::printf("type,vec_entries,vec_MB,S,U,gathers,avg_us,Ggathers_s,GB_s_useful\n");
    for (int64_t n : sizes) {
        for (int S : splits) {
            // Workgroup w of the U = 1 launch owns indices [w*256, (w+1)*256); for U > 1 the
            // workgroup owns U such groups, so the block of index q is (q / 256) / U mod S.
            // Blocks depend on U; regenerate for each U so that every launch obeys the split.
            for (int U : {1, 4, 8, 16}) {
                const int64_t blk = n / S;
                for (int64_t q = 0; q < gathers; ++q) {
                    const int64_t w = (q / kThreads) / U;
                    const int64_t k = w % S;
                    idx[q] = static_cast<int>(k * blk + static_cast<int64_t>(next_u64() % static_cast<uint64_t>(blk)));
                    if (idx[q] < 0 || idx[q] >= n) {
                        std::fprintf(stderr, "index out of range\n");
                        return 3;
                    }
                }
                CK(hipMemcpy(d_idx, idx.data(), gathers * sizeof(int), hipMemcpyHostToDevice));
                auto one = [&](auto tag, const char* name) {
                    using T = decltype(tag);
                    if (n * static_cast<int64_t>(sizeof(T)) > nmax * static_cast<int64_t>(sizeof(double))) return;
                    T* v = reinterpret_cast<T*>(d_vec);
                    T* o = reinterpret_cast<T*>(d_out);
                    double us = 0.0;
                    switch (U) {
                        case 1: us = time_kernel<T, 1>(gathers, d_idx, v, o, reps, s); break;
                        case 4: us = time_kernel<T, 4>(gathers, d_idx, v, o, reps, s); break;
                        case 8: us = time_kernel<T, 8>(gathers, d_idx, v, o, reps, s); break;
                        default: us = time_kernel<T, 16>(gathers, d_idx, v, o, reps, s); break;
                    }
                    const double g = static_cast<double>(gathers) / (us * 1e-6) / 1e9;
                    std::printf("%s,%lld,%.1f,%d,%d,%lld,%.2f,%.1f,%.1f\n", name, static_cast<long long>(n),
                                n * sizeof(T) / 1e6, S, U, static_cast<long long>(gathers), us, g, g * sizeof(T));
                    std::fflush(stdout);
                };
                one(float{}, "f32");
                one(double{}, "f64");
            }
        }
    }
    CK(hipFree(d_vec));
    CK(hipFree(d_out));
    CK(hipFree(d_idx));
    CK(hipStreamDestroy(s));
    return 0;
}
