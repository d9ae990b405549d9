// Device-side building blocks for the CDNA4 kernels: value loads for FP64 / FP32 / BF16
// matrices, reductions across a lane group of a 64-wide wavefront, and deterministic
// per-tile (per-instance) partial sums.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

namespace vj {

constexpr int kBlock = 256;          // threads per workgroup (4 wavefronts of 64)
constexpr int kWave = 64;
constexpr int kEllSlice = 64;        // rows per slice of the sliced-ELL layout (one wavefront)

// BF16 matrix values are stored as raw 16-bit patterns (upper half of an FP32); FP16 values
// as IEEE binary16 patterns.
struct bf16_t { uint16_t bits; };
struct f16_t { uint16_t bits; };

__device__ __forceinline__ float f16_bits_to_float(uint16_t b) { return static_cast<float>(__builtin_bit_cast(_Float16, b)); }
// Round to nearest even (values beyond the FP16 range become infinite; the engine checks the
// range before it converts).
__device__ __forceinline__ uint16_t f16_bits_from(double v) {
    return __builtin_bit_cast(uint16_t, static_cast<_Float16>(static_cast<float>(v)));
}

// Matrix streams are read once per iteration: non-temporal loads keep them from evicting
// the gathered iterate vectors from L2 and the Infinity Cache.
__device__ __forceinline__ double load_val(const double* p, int64_t k) { return __builtin_nontemporal_load(p + k); }
__device__ __forceinline__ float load_val(const float* p, int64_t k) { return __builtin_nontemporal_load(p + k); }
__device__ __forceinline__ float load_val(const bf16_t* p, int64_t k) {
    const uint16_t b = __builtin_nontemporal_load(&p[k].bits);
    return __uint_as_float(static_cast<uint32_t>(b) << 16);
}
__device__ __forceinline__ float load_val(const f16_t* p, int64_t k) {
    return f16_bits_to_float(__builtin_nontemporal_load(&p[k].bits));
}
__device__ __forceinline__ int load_idx(const int* p, int64_t k) { return __builtin_nontemporal_load(p + k); }

// Sum over the G consecutive lanes of a group (G a power of two, G <= 64). Fixed
// butterfly order, so the result is deterministic; every lane receives the sum.
template <int G, class T>
__device__ __forceinline__ T group_sum(T v) {
#pragma unroll
    for (int off = G / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, G);
    return v;
}
template <int G, class T>
__device__ __forceinline__ T group_max(T v) {
#pragma unroll
    for (int off = G / 2; off > 0; off >>= 1) {
        T o = __shfl_xor(v, off, G);
        v = (v < o) ? o : v;
    }
    return v;
}

// Block-wide sum of NQ doubles (256 threads). Result valid in thread 0.
template <int NQ>
__device__ __forceinline__ void block_sum(double (&v)[NQ], double (*lds)[NQ]) {
#pragma unroll
    for (int q = 0; q < NQ; ++q) v[q] = group_sum<kWave>(v[q]);
    const int wave = threadIdx.x / kWave;
    if ((threadIdx.x & (kWave - 1)) == 0) {
#pragma unroll
        for (int q = 0; q < NQ; ++q) lds[wave][q] = v[q];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            double s = lds[0][q];
            for (int w = 1; w < kBlock / kWave; ++w) s += lds[w][q];
            v[q] = s;
        }
    }
    __syncthreads();
}

// Write the partial sums of a tile of rows [r0, r0+R) to part[(tile + inst) * NQ + q]
// for every instance overlapping the tile (rows are grouped by instance, so the slot
// index tile+inst is unique and increases with both). `my_inst` is the instance of the
// row whose contribution this thread holds (ignored when MULTI is false). Column q is
// written only when bit q of `mask` is set.
template <int NQ, bool MULTI>
__device__ __forceinline__ void write_tile_partials(const double (&v)[NQ], int my_inst,
                                                    int inst_first, int inst_last, int tile,
                                                    double* __restrict__ part, int mask) {
    __shared__ double lds[kBlock / kWave][NQ];
    if (!MULTI) {
        double t[NQ];
#pragma unroll
        for (int q = 0; q < NQ; ++q) t[q] = v[q];
        block_sum<NQ>(t, lds);
        if (threadIdx.x == 0) {
#pragma unroll
            for (int q = 0; q < NQ; ++q)
                if ((mask >> q) & 1) part[static_cast<int64_t>(tile) * NQ + q] = t[q];
        }
        return;
    }
    for (int inst = inst_first; inst <= inst_last; ++inst) {
        double t[NQ];
#pragma unroll
        for (int q = 0; q < NQ; ++q) t[q] = (my_inst == inst) ? v[q] : 0.0;
        block_sum<NQ>(t, lds);
        if (threadIdx.x == 0) {
#pragma unroll
            for (int q = 0; q < NQ; ++q)
                if ((mask >> q) & 1) part[static_cast<int64_t>(tile + inst) * NQ + q] = t[q];
        }
    }
}

// Deterministic sum, over the tiles of one instance, of partial column q.
// Called by all threads of a 256-thread block; result in thread 0.
__device__ __forceinline__ double reduce_instance(const double* __restrict__ part, int NQ, int q,
                                                  int inst, int64_t first_row, int64_t end_row,
                                                  int R, double* lds256) {
    double s = 0.0;
    if (end_row > first_row) {
        const int t0 = static_cast<int>(first_row / R);
        const int t1 = static_cast<int>((end_row - 1) / R);
        for (int t = t0 + static_cast<int>(threadIdx.x); t <= t1; t += kBlock)
            s += part[static_cast<int64_t>(t + inst) * NQ + q];
    }
    lds256[threadIdx.x] = s;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if (static_cast<int>(threadIdx.x) < w) lds256[threadIdx.x] += lds256[threadIdx.x + w];
        __syncthreads();
    }
    const double r = lds256[0];
    __syncthreads();
    return r;
}

// Same over an explicit tile range [t0, t1] of the instance (variable-height tiles).
__device__ __forceinline__ double reduce_tiles(const double* __restrict__ part, int NQ, int q, int inst, int t0,
                                               int t1, double* lds256) {
    double s = 0.0;
    for (int t = t0 + static_cast<int>(threadIdx.x); t <= t1; t += kBlock)
        s += part[static_cast<int64_t>(t + inst) * NQ + q];
    lds256[threadIdx.x] = s;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if (static_cast<int>(threadIdx.x) < w) lds256[threadIdx.x] += lds256[threadIdx.x + w];
        __syncthreads();
    }
    const double r = lds256[0];
    __syncthreads();
    return r;
}

}  // namespace vj
