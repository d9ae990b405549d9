// Device-side construction of the matrix layouts of the restarted engine: the CSR of A from
// its CSC (the reference Model layout; replaces SparseMatrixCSC::to_csr,
// src/linalg/sparse.cpp:147-184, which the reference runs on the host before its GPU solve,
// pdhg_cuda.cu:159), the row tiles of the tile-streamed kernels, the XCD column-split copy
// and the sliced-ELL copy. Every layout is a pure function of the sparsity pattern and is
// built deterministically (no order depends on scheduling), so it is the same on every run.
#include <hip/hip_runtime.h>

#include <cstring>
#include <vector>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include "vj_device.hpp"
#include "vj_layout.hpp"

namespace vj {

namespace {

// src_row[p] = the input row holding entry p (the last r with ptr[r] <= p).
__global__ void entry_rows(int64_t nnz, int rows, const int* __restrict__ ptr, int* __restrict__ src_row) {
    const int64_t p = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (p >= nnz) return;
    int lo = 0, hi = rows;   // ptr[lo] <= p < ptr[hi]
    while (hi - lo > 1) {
        const int mid = lo + (hi - lo) / 2;
        if (ptr[mid] <= p) lo = mid; else hi = mid;
    }
    src_row[p] = lo;
}

__global__ void iota(int64_t nnz, int* __restrict__ v) {
    const int64_t p = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (p < nnz) v[p] = static_cast<int>(p);
}

// Output row pointers from the sorted keys: out_ptr[r] = first q with key[q] >= r.
__global__ void ptr_from_sorted(int rows, int64_t nnz, const int* __restrict__ key, int* __restrict__ out_ptr) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r > rows) return;
    int64_t lo = 0, hi = nnz;
    while (lo < hi) {
        const int64_t mid = lo + (hi - lo) / 2;
        if (key[mid] < r) lo = mid + 1; else hi = mid;
    }
    out_ptr[r] = static_cast<int>(lo);
}

__global__ void gather_transposed(int64_t nnz, const int* __restrict__ perm, const int* __restrict__ src_row,
                                  const double* __restrict__ in_val, int* __restrict__ out_idx,
                                  double* __restrict__ out_val) {
    const int64_t q = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (q >= nnz) return;
    const int p = perm[q];
    out_idx[q] = src_row[p];
    out_val[q] = in_val[p];
}

// next[r]: the end of a tile starting at row r: the largest e in [r+1, min(rows, r+max_rows)]
// with ptr_k[e] - ptr_k[r] <= cap in every one of the S pointer arrays (ptr_k = ptrs +
// k*stride), or r+1 when row r alone exceeds cap in some array (a tile of its own). With
// inst_of set, a tile also ends where the instance of row r ends (off[inst_of[r] + 1]).
__global__ void tile_next(int rows, int S, const int* __restrict__ ptrs, int64_t stride, int cap, int max_rows,
                          const int* __restrict__ inst_of, const int* __restrict__ off, int* __restrict__ next) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= rows) return;
    auto fits = [&](int e) {
        for (int k = 0; k < S; ++k) {
            const int* p = ptrs + k * stride;
            if (p[e] - p[r] > cap) return false;
        }
        return true;
    };
    int lo = r + 1, hi = r + max_rows < rows ? r + max_rows : rows;   // answer in [lo, hi]
    if (inst_of) {
        const int end = off[inst_of[r] + 1];
        hi = end < hi ? end : hi;
    }
    if (!fits(lo)) {
        next[r] = r + 1;
        return;
    }
    while (lo < hi) {   // fits(lo) holds
        const int mid = lo + (hi - lo + 1) / 2;
        if (fits(mid)) lo = mid; else hi = mid - 1;
    }
    next[r] = lo;
}

// One thread follows the chain 0 -> next[0] -> ... -> rows: tile_row[t] is the first row of
// tile t, tile_row[count] = rows; count goes to *ntiles.
__global__ void tile_chain(int rows, const int* __restrict__ next, int* __restrict__ tile_row, int* __restrict__ ntiles) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    int t = 0, r = 0;
    tile_row[0] = 0;
    while (r < rows) {
        r = next[r];
        tile_row[++t] = r;
    }
    *ntiles = t;
}

// first[q] / last[q]: the tiles holding the first and the last row of instance q.
__global__ void instance_tiles(int count, int ntiles, const int* __restrict__ tile_row, const int* __restrict__ off,
                               int* __restrict__ first, int* __restrict__ last) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= count) return;
    auto tile_of = [&](int row) {   // the last t with tile_row[t] <= row
        int lo = 0, hi = ntiles;
        while (hi - lo > 1) {
            const int mid = lo + (hi - lo) / 2;
            if (tile_row[mid] <= row) lo = mid; else hi = mid;
        }
        return lo;
    };
    first[q] = tile_of(off[q]);
    last[q] = tile_of(off[q + 1] - 1);
}

// win[2t], win[2t+1]: the least column index of tile t's entries and one past the greatest
// (0, 0 for a tile without entries). One workgroup per tile.
__global__ __launch_bounds__(kBlock) void tile_windows(int ntiles, const int* __restrict__ tile_row,
                                                       const int* __restrict__ ptr, const int* __restrict__ idx,
                                                       int* __restrict__ win) {
    __shared__ int lo_s[kBlock / kWave], hi_s[kBlock / kWave];
    const int t = blockIdx.x;
    if (t >= ntiles) return;
    const int s0 = ptr[tile_row[t]], s1 = ptr[tile_row[t + 1]];
    int lo = 2147483647, hi = -1;
    for (int p = s0 + static_cast<int>(threadIdx.x); p < s1; p += kBlock) {
        const int c = idx[p];
        lo = c < lo ? c : lo;
        hi = c > hi ? c : hi;
    }
    for (int o = kWave / 2; o > 0; o >>= 1) {
        const int a = __shfl_xor(lo, o, kWave), b = __shfl_xor(hi, o, kWave);
        lo = a < lo ? a : lo;
        hi = b > hi ? b : hi;
    }
    const int w = threadIdx.x / kWave;
    if ((threadIdx.x & (kWave - 1)) == 0) {
        lo_s[w] = lo;
        hi_s[w] = hi;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < kBlock / kWave; ++k) {
            lo = lo_s[k] < lo ? lo_s[k] : lo;
            hi = hi_s[k] > hi ? hi_s[k] : hi;
        }
        win[2 * t] = hi < 0 ? 0 : lo;
        win[2 * t + 1] = hi < 0 ? 0 : hi + 1;
    }
}

// off16[p] = idx[p] - win[2t] for the entries p of tile t (one workgroup per tile; every
// window is at most 65,536 columns wide).
__global__ __launch_bounds__(kBlock) void window_offsets(int ntiles, const int* __restrict__ tile_row,
                                                         const int* __restrict__ ptr, const int* __restrict__ idx,
                                                         const int* __restrict__ win, uint16_t* __restrict__ off16) {
    const int t = blockIdx.x;
    if (t >= ntiles) return;
    const int s0 = ptr[tile_row[t]], s1 = ptr[tile_row[t + 1]];
    const int lo = win[2 * t];
    for (int p = s0 + static_cast<int>(threadIdx.x); p < s1; p += kBlock) off16[p] = static_cast<uint16_t>(idx[p] - lo);
}

// desc[8t .. 8t+7] = {first row, end row, first entry, end entry, window start, window end,
// instance, 0} of tile t: everything a staged workgroup needs before its loads, in one place.
__global__ void stage_descs(int ntiles, const int* __restrict__ tile_row, const int* __restrict__ ptr,
                            const int* __restrict__ win, const int* __restrict__ inst, int* __restrict__ desc) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntiles) return;
    const int r0 = tile_row[t], r1 = tile_row[t + 1];
    int* d = desc + 8 * static_cast<int64_t>(t);
    d[0] = r0;
    d[1] = r1;
    d[2] = ptr[r0];
    d[3] = ptr[r1];
    d[4] = win[2 * t];
    d[5] = win[2 * t + 1];
    d[6] = inst[t];
    d[7] = 0;
}

// inst[t] = the instance of tile t's first row (0 without an instance map).
__global__ void tile_instances(int ntiles, const int* __restrict__ tile_row, const int* __restrict__ inst_of,
                               int* __restrict__ inst) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < ntiles) inst[t] = inst_of ? inst_of[tile_row[t]] : 0;
}

__global__ void tile_bases(int ntiles, int S, const int* __restrict__ tile_row, const int* __restrict__ ptrs,
                           int64_t stride, int* __restrict__ tbase) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= static_cast<int64_t>(ntiles + 1) * S) return;
    const int t = static_cast<int>(i / S), k = static_cast<int>(i % S);
    tbase[i] = ptrs[k * stride + tile_row[t]];
}

// toff[i] for row r = i mod rows of block k = i / rows: its first entry in the block minus the
// first entry of its tile in the block.
__global__ void tile_offsets(int rows, int S, int ntiles, const int* __restrict__ tile_row, const int* __restrict__ ptrs,
                             int64_t stride, uint16_t* __restrict__ toff) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= static_cast<int64_t>(S) * rows) return;
    const int64_t k = i / rows;
    const int r = static_cast<int>(i % rows);
    int lo = 0, hi = ntiles;   // the last tile t with tile_row[t] <= r
    while (hi - lo > 1) {
        const int mid = lo + (hi - lo) / 2;
        if (tile_row[mid] <= r) lo = mid; else hi = mid;
    }
    const int v = ptrs[k * stride + r] - ptrs[k * stride + tile_row[lo]];
    toff[i] = static_cast<uint16_t>(v > 65535 ? 65535 : v);
}

// Column split: F[k*(rows+1) + r + 1] = entries of row r in column block k, F[k*(rows+1)] = 0;
// G lanes per row.
constexpr int kSplitLanes = 8;
__global__ __launch_bounds__(kBlock) void split_counts(int rows, int S, int64_t B, const int* __restrict__ rp,
                                                       const int* __restrict__ ci, int* __restrict__ F) {
    const int r = blockIdx.x * (kBlock / kSplitLanes) + static_cast<int>(threadIdx.x) / kSplitLanes;
    const int lane = threadIdx.x % kSplitLanes;
    int cnt[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (r < rows)
        for (int p = rp[r] + lane; p < rp[r + 1]; p += kSplitLanes) {
            const int k = static_cast<int>(ci[p] / B);
#pragma unroll
            for (int u = 0; u < 8; ++u) cnt[u] += (k == u);
        }
#pragma unroll
    for (int u = 0; u < 8; ++u) cnt[u] = group_sum<kSplitLanes>(cnt[u]);
    if (r < rows && lane == 0) {
        const int64_t stride = static_cast<int64_t>(rows) + 1;
        for (int k = 0; k < S; ++k) {
            F[k * stride + r + 1] = cnt[k];
            if (r == 0) F[k * stride] = 0;
        }
    }
}

// Split scatter: the entries of row r go to their blocks in their CSR order. A group of G
// lanes walks the row in chunks of G entries; within a chunk, a lane's slot in block k is
// the number of earlier lanes of its group whose entry is in block k (ballot + popcount).
__global__ __launch_bounds__(kBlock) void split_scatter(int rows, int S, int64_t B, const int* __restrict__ rp,
                                                        const int* __restrict__ ci, const int* __restrict__ bp,
                                                        int* __restrict__ idx, int* __restrict__ perm) {
    const int r = blockIdx.x * (kBlock / kSplitLanes) + static_cast<int>(threadIdx.x) / kSplitLanes;
    const int lane = threadIdx.x % kSplitLanes;
    const int wlane = threadIdx.x % kWave;
    const int gshift = wlane - lane;   // first bit of this group in the wave's ballot
    const uint64_t gmask = ((uint64_t{1} << kSplitLanes) - 1) << gshift;
    const uint64_t below = (uint64_t{1} << wlane) - 1;
    const int64_t stride = static_cast<int64_t>(rows) + 1;
    int cur[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    int p0 = 0, p1 = 0;
    if (r < rows) {
        p0 = rp[r];
        p1 = rp[r + 1];
#pragma unroll
        for (int k = 0; k < 8; ++k)
            if (k < S) cur[k] = bp[k * stride + r];
    }
    // every lane of the wave takes part in each ballot: the loop runs to the longest row of
    // the wave's groups
    int len = p1 - p0;
    for (int off = kWave / 2; off > 0; off >>= 1) {
        const int o = __shfl_xor(len, off, kWave);
        len = len > o ? len : o;
    }
    for (int c = 0; c < len; c += kSplitLanes) {
        const int p = p0 + c + lane;
        const bool mine = r < rows && p < p1;
        const int col = mine ? ci[p] : 0;
        const int kb = mine ? static_cast<int>(col / B) : -1;
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const uint64_t m = __ballot(kb == k) & gmask;
            if (kb == k) {
                const int q = cur[k] + __popcll(m & below);
                idx[q] = col;
                perm[q] = p;
            }
            cur[k] += __popcll(m);
        }
    }
}

// Sliced ELL: the width of each slice of kEllSlice rows (its longest row).
__global__ void ell_widths(int rows, int slices, const int* __restrict__ rp, int* __restrict__ width) {
    const int sl = blockIdx.x * blockDim.x + threadIdx.x;
    if (sl >= slices) return;
    int w = 0;
    const int r1 = (sl + 1) * kEllSlice < rows ? (sl + 1) * kEllSlice : rows;
    for (int r = sl * kEllSlice; r < r1; ++r) w = max(w, rp[r + 1] - rp[r]);
    width[sl] = w;
}

__global__ void ell_fill(int rows, const int* __restrict__ rp, const int* __restrict__ ci,
                         const int64_t* __restrict__ off, const int* __restrict__ width, int* __restrict__ idx,
                         int* __restrict__ perm) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    const int sl = r / kEllSlice;
    if (sl * kEllSlice >= rows) return;
    const int lane = r % kEllSlice;
    const int len = r < rows ? rp[r + 1] - rp[r] : 0;
    const int first = len > 0 ? ci[rp[r]] : 0;
    const int w = width[sl];
    for (int q = 0; q < w; ++q) {
        const int64_t e = off[sl] + static_cast<int64_t>(q) * kEllSlice + lane;
        if (q < len) {
            idx[e] = ci[rp[r] + q];
            perm[e] = rp[r] + q;
        } else {
            idx[e] = first;
            perm[e] = -1;
        }
    }
}

__global__ void narrow_offsets(int slices, const int64_t* __restrict__ off, int* __restrict__ off32) {
    const int sl = blockIdx.x * blockDim.x + threadIdx.x;
    if (sl < slices) off32[sl] = static_cast<int>(off[sl]);
}

__global__ void scale_widths(int slices, const int* __restrict__ w, int64_t* __restrict__ sz) {
    const int sl = blockIdx.x * blockDim.x + threadIdx.x;
    if (sl < slices) sz[sl] = static_cast<int64_t>(w[sl]) * kEllSlice;
}

template <class T>
int exclusive_scan(hipStream_t s, const T* in, T* out, size_t len, DevBuf& tmp) {
    size_t bytes = 0;
    VJ_HIP(rocprim::exclusive_scan(nullptr, bytes, in, out, T(0), len, rocprim::plus<T>(), s));
    if (tmp.bytes < bytes) VJ_TRY(tmp.alloc(bytes));
    VJ_HIP(rocprim::exclusive_scan(tmp.p, bytes, in, out, T(0), len, rocprim::plus<T>(), s));
    return VJ_OK;
}

template <class T>
int inclusive_scan(hipStream_t s, const T* in, T* out, size_t len, DevBuf& tmp) {
    size_t bytes = 0;
    VJ_HIP(rocprim::inclusive_scan(nullptr, bytes, in, out, len, rocprim::plus<T>(), s));
    if (tmp.bytes < bytes) VJ_TRY(tmp.alloc(bytes));
    VJ_HIP(rocprim::inclusive_scan(tmp.p, bytes, in, out, len, rocprim::plus<T>(), s));
    return VJ_OK;
}

unsigned key_bits(int bound) {
    unsigned b = 1;
    while (b < 31 && (1u << b) < static_cast<unsigned>(bound)) ++b;
    return b;
}

}  // namespace

int device_transpose(hipStream_t s, int in_rows, int out_rows, int64_t nnz, const int* in_ptr, const int* in_idx,
                     const double* in_val, DevBuf& out_ptr, DevBuf& out_idx, DevBuf& out_val) {
    DevBuf src_row, keys, perm0, perm, tmp;
    VJ_TRY(src_row.alloc(static_cast<size_t>(nnz) * 4));
    VJ_TRY(keys.alloc(static_cast<size_t>(nnz) * 4));
    VJ_TRY(perm0.alloc(static_cast<size_t>(nnz) * 4));
    VJ_TRY(perm.alloc(static_cast<size_t>(nnz) * 4));
    VJ_TRY(out_ptr.alloc((static_cast<size_t>(out_rows) + 1) * 4));
    VJ_TRY(out_idx.alloc(static_cast<size_t>(nnz) * 4));
    VJ_TRY(out_val.alloc(static_cast<size_t>(nnz) * 8));
    const int g = grid_for(nnz, kBlock);
    hipLaunchKernelGGL(entry_rows, dim3(g), dim3(kBlock), 0, s, nnz, in_rows, in_ptr, src_row.as<int>());
    hipLaunchKernelGGL(iota, dim3(g), dim3(kBlock), 0, s, nnz, perm0.as<int>());
    VJ_LAUNCHED();
    // stable (LSD radix) sort of the entries by output row: entries of one output row keep
    // their input order, i.e. ascending input row, the column order to_csr produces
    size_t bytes = 0;
    const unsigned bits = key_bits(out_rows);
    const unsigned* kin = reinterpret_cast<const unsigned*>(in_idx);   // indices are >= 0
    VJ_HIP(rocprim::radix_sort_pairs(nullptr, bytes, kin, keys.as<unsigned>(), perm0.as<int>(), perm.as<int>(),
                                     static_cast<size_t>(nnz), 0, bits, s));
    VJ_TRY(tmp.alloc(bytes));
    VJ_HIP(rocprim::radix_sort_pairs(tmp.p, bytes, kin, keys.as<unsigned>(), perm0.as<int>(), perm.as<int>(),
                                     static_cast<size_t>(nnz), 0, bits, s));
    hipLaunchKernelGGL(ptr_from_sorted, dim3(grid_for(static_cast<int64_t>(out_rows) + 1, kBlock)), dim3(kBlock), 0, s,
                       out_rows, nnz, keys.as<int>(), out_ptr.as<int>());
    hipLaunchKernelGGL(gather_transposed, dim3(g), dim3(kBlock), 0, s, nnz, perm.as<int>(), src_row.as<int>(), in_val,
                       out_idx.as<int>(), out_val.as<double>());
    VJ_LAUNCHED();
    VJ_HIP(hipStreamSynchronize(s));   // the scratch buffers of this call are released on return
    return VJ_OK;
}

int device_row_tiles(hipStream_t s, int rows, int S, const int* ptrs, int64_t stride, int cap, int max_rows,
                     DevBuf& tile_row, int* ntiles, const int* inst_of, const int* off) {
    DevBuf next, tmp, cnt;
    VJ_TRY(next.alloc(static_cast<size_t>(rows) * 4));
    VJ_TRY(tmp.alloc((static_cast<size_t>(rows) + 1) * 4));
    VJ_TRY(cnt.alloc(4));
    hipLaunchKernelGGL(tile_next, dim3(grid_for(rows, kBlock)), dim3(kBlock), 0, s, rows, S, ptrs, stride, cap, max_rows,
                       inst_of, off, next.as<int>());
    hipLaunchKernelGGL(tile_chain, dim3(1), dim3(1), 0, s, rows, next.as<int>(), tmp.as<int>(), cnt.as<int>());
    VJ_LAUNCHED();
    int t = 0;
    VJ_HIP(hipMemcpyAsync(&t, cnt.p, 4, hipMemcpyDeviceToHost, s));
    VJ_HIP(hipStreamSynchronize(s));
    VJ_TRY(tile_row.alloc((static_cast<size_t>(t) + 1) * 4));
    VJ_HIP(hipMemcpyAsync(tile_row.p, tmp.p, (static_cast<size_t>(t) + 1) * 4, hipMemcpyDeviceToDevice, s));
    VJ_HIP(hipStreamSynchronize(s));
    *ntiles = t;
    return VJ_OK;
}

int device_tile_windows(hipStream_t s, int ntiles, const int* tile_row, const int* ptr, const int* idx, DevBuf& win,
                        int64_t* total, int* widest) {
    VJ_TRY(win.alloc(std::max<size_t>(static_cast<size_t>(ntiles), 1) * 2 * 4));
    if (ntiles > 0) {
        hipLaunchKernelGGL(tile_windows, dim3(ntiles), dim3(kBlock), 0, s, ntiles, tile_row, ptr, idx, win.as<int>());
        VJ_LAUNCHED();
    }
    std::vector<int> h(static_cast<size_t>(ntiles) * 2);
    if (ntiles > 0) VJ_HIP(hipMemcpyAsync(h.data(), win.p, h.size() * 4, hipMemcpyDeviceToHost, s));
    VJ_HIP(hipStreamSynchronize(s));
    int64_t sum = 0;
    int w = 0;
    for (int t = 0; t < ntiles; ++t) {
        sum += h[2 * t + 1] - h[2 * t];
        w = std::max(w, h[2 * t + 1] - h[2 * t]);
    }
    *total = sum;
    *widest = w;
    return VJ_OK;
}

int device_window_offsets(hipStream_t s, int ntiles, const int* tile_row, const int* ptr, const int* idx,
                          const int* win, int64_t nnz, DevBuf& off16) {
    VJ_TRY(off16.alloc(std::max<size_t>(static_cast<size_t>(nnz), 1) * 2));
    if (ntiles > 0) {
        hipLaunchKernelGGL(window_offsets, dim3(ntiles), dim3(kBlock), 0, s, ntiles, tile_row, ptr, idx, win,
                           off16.as<uint16_t>());
        VJ_LAUNCHED();
    }
    return VJ_OK;
}

int device_tile_instances(hipStream_t s, int ntiles, const int* tile_row, const int* inst_of, DevBuf& inst) {
    VJ_TRY(inst.alloc(std::max<size_t>(static_cast<size_t>(ntiles), 1) * 4));
    if (ntiles > 0) {
        hipLaunchKernelGGL(tile_instances, dim3(grid_for(ntiles, kBlock)), dim3(kBlock), 0, s, ntiles, tile_row, inst_of,
                           inst.as<int>());
        VJ_LAUNCHED();
    }
    return VJ_OK;
}

int device_stage_descs(hipStream_t s, int ntiles, const int* tile_row, const int* ptr, const int* win,
                       const int* inst, DevBuf& desc) {
    VJ_TRY(desc.alloc(std::max<size_t>(static_cast<size_t>(ntiles), 1) * 32));
    if (ntiles > 0) {
        hipLaunchKernelGGL(stage_descs, dim3(grid_for(ntiles, kBlock)), dim3(kBlock), 0, s, ntiles, tile_row, ptr, win,
                           inst, desc.as<int>());
        VJ_LAUNCHED();
    }
    return VJ_OK;
}

int device_instance_tiles(hipStream_t s, int count, int ntiles, const int* tile_row, const int* off, DevBuf& first,
                          DevBuf& last) {
    VJ_TRY(first.alloc(static_cast<size_t>(count) * 4));
    VJ_TRY(last.alloc(static_cast<size_t>(count) * 4));
    hipLaunchKernelGGL(instance_tiles, dim3(grid_for(count, kBlock)), dim3(kBlock), 0, s, count, ntiles, tile_row, off,
                       first.as<int>(), last.as<int>());
    VJ_LAUNCHED();
    return VJ_OK;
}

int device_tile_bases(hipStream_t s, int ntiles, int S, const int* tile_row, const int* ptrs, int64_t stride,
                      DevBuf& tbase) {
    const int64_t len = static_cast<int64_t>(ntiles + 1) * S;
    VJ_TRY(tbase.alloc(static_cast<size_t>(len) * 4));
    hipLaunchKernelGGL(tile_bases, dim3(grid_for(len, kBlock)), dim3(kBlock), 0, s, ntiles, S, tile_row, ptrs, stride,
                       tbase.as<int>());
    VJ_LAUNCHED();
    return VJ_OK;
}

int device_tile_offsets(hipStream_t s, int rows, int S, int ntiles, const int* tile_row, const int* ptrs,
                        int64_t stride, DevBuf& toff) {
    const int64_t len = static_cast<int64_t>(S) * rows;
    VJ_TRY(toff.alloc(std::max<size_t>(static_cast<size_t>(len), 1) * 2));
    hipLaunchKernelGGL(tile_offsets, dim3(grid_for(len, kBlock)), dim3(kBlock), 0, s, rows, S, ntiles, tile_row, ptrs,
                       stride, toff.as<uint16_t>());
    VJ_LAUNCHED();
    return VJ_OK;
}

int device_split(hipStream_t s, int rows, int cols, int64_t nnz, const int* rp, const int* ci, int S, DevBuf& bp,
                 DevBuf& idx, DevBuf& perm) {
    if (S < 2 || S > 8) {
        set_error("column split into " + std::to_string(S) + " blocks (2..8 supported)");
        return VJ_ERR_INVALID_ARGUMENT;
    }
    const int64_t B = (static_cast<int64_t>(cols) + S - 1) / S;
    const size_t len = static_cast<size_t>(S) * (static_cast<size_t>(rows) + 1);
    DevBuf F, tmp;
    VJ_TRY(F.alloc(len * 4));
    VJ_TRY(bp.alloc(len * 4));
    VJ_TRY(idx.alloc(static_cast<size_t>(nnz) * 4));
    VJ_TRY(perm.alloc(static_cast<size_t>(nnz) * 4));
    const int g = grid_for(rows, kBlock / kSplitLanes);
    hipLaunchKernelGGL(split_counts, dim3(g), dim3(kBlock), 0, s, rows, S, B, rp, ci, F.as<int>());
    VJ_LAUNCHED();
    // bp[k][r] = entries of blocks before k + entries of block k in rows before r
    VJ_TRY(inclusive_scan<int>(s, F.as<int>(), bp.as<int>(), len, tmp));
    hipLaunchKernelGGL(split_scatter, dim3(g), dim3(kBlock), 0, s, rows, S, B, rp, ci, bp.as<int>(), idx.as<int>(),
                       perm.as<int>());
    VJ_LAUNCHED();
    VJ_HIP(hipStreamSynchronize(s));
    return VJ_OK;
}

int device_ell(hipStream_t s, int rows, int64_t nnz, const int* rp, const int* ci, double max_padding, bool forced,
               double* padding, int64_t* ell_len, DevBuf& off32, DevBuf& width, DevBuf& idx, DevBuf& perm) {
    const int slices = (rows + kEllSlice - 1) / kEllSlice;
    DevBuf sz, off, tmp;
    VJ_TRY(width.alloc(static_cast<size_t>(slices) * 4));
    VJ_TRY(sz.alloc((static_cast<size_t>(slices) + 1) * 8));
    VJ_TRY(off.alloc((static_cast<size_t>(slices) + 1) * 8));
    hipLaunchKernelGGL(ell_widths, dim3(grid_for(slices, kBlock)), dim3(kBlock), 0, s, rows, slices, rp, width.as<int>());
    hipLaunchKernelGGL(scale_widths, dim3(grid_for(slices, kBlock)), dim3(kBlock), 0, s, slices, width.as<int>(),
                       sz.as<int64_t>());
    VJ_LAUNCHED();
    VJ_HIP(hipMemsetAsync(sz.as<int64_t>() + slices, 0, 8, s));
    VJ_TRY(exclusive_scan<int64_t>(s, sz.as<int64_t>(), off.as<int64_t>(), static_cast<size_t>(slices) + 1, tmp));
    int64_t total = 0;
    VJ_HIP(hipMemcpyAsync(&total, off.as<int64_t>() + slices, 8, hipMemcpyDeviceToHost, s));
    VJ_HIP(hipStreamSynchronize(s));
    *padding = nnz > 0 ? static_cast<double>(total) / static_cast<double>(nnz) : 0.0;
    *ell_len = 0;
    if (total > 2147483647LL) {   // beyond 32-bit offsets: no ELL copy
        *padding = 0.0;
        return VJ_OK;
    }
    if (*padding > max_padding && !forced) return VJ_OK;
    VJ_TRY(off32.alloc(static_cast<size_t>(slices) * 4));
    VJ_TRY(idx.alloc(std::max<size_t>(static_cast<size_t>(total), 1) * 4));
    VJ_TRY(perm.alloc(std::max<size_t>(static_cast<size_t>(total), 1) * 4));
    hipLaunchKernelGGL(ell_fill, dim3(grid_for(static_cast<int64_t>(slices) * kEllSlice, kBlock)), dim3(kBlock), 0, s, rows,
                       rp, ci, off.as<int64_t>(), width.as<int>(), idx.as<int>(), perm.as<int>());
    hipLaunchKernelGGL(narrow_offsets, dim3(grid_for(slices, kBlock)), dim3(kBlock), 0, s, slices, off.as<int64_t>(),
                       off32.as<int>());
    VJ_LAUNCHED();
    VJ_HIP(hipStreamSynchronize(s));
    *ell_len = total;
    return VJ_OK;
}

}  // namespace vj
