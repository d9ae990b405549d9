// Device-side layout builders of the restarted engine (vj_layout.hip). Every function works
// on device arrays on stream `s` and returns a VJ_* status; the layouts are deterministic.
#pragma once
#include "vj_internal.hpp"

namespace vj {

// CSR of B^T from the CSR of B (in_rows rows, in_ptr int32 [in_rows+1], in_idx < out_rows):
// out_ptr [out_rows+1], out_idx, out_val; within each output row the entries ascend by input
// row, as SparseMatrixCSC::to_csr orders them.
int device_transpose(hipStream_t s, int in_rows, int out_rows, int64_t nnz, const int* in_ptr, const int* in_idx,
                     const double* in_val, DevBuf& out_ptr, DevBuf& out_idx, DevBuf& out_val);

// Row tiles: consecutive rows, at most max_rows rows and at most cap entries in each of the S
// pointer arrays ptrs + k*stride (k < S); a row over cap in some array is a tile of its own.
// With inst_of / off (row -> instance, instance row offsets) no tile spans two instances.
// tile_row [ntiles+1] holds the first row of each tile and rows at the end.
int device_row_tiles(hipStream_t s, int rows, int S, const int* ptrs, int64_t stride, int cap, int max_rows,
                     DevBuf& tile_row, int* ntiles, const int* inst_of = nullptr, const int* off = nullptr);

// Column windows of row tiles of a CSR (ptr, idx): win [2*ntiles], tile t's entries read
// columns in [win[2t], win[2t+1]) (0, 0 for a tile without entries); *total = the sum of the
// window widths, *widest = the widest window.
int device_tile_windows(hipStream_t s, int ntiles, const int* tile_row, const int* ptr, const int* idx, DevBuf& win,
                        int64_t* total, int* widest);

// 16-bit window offsets of the entries: off16[p] = idx[p] - win[2t] for every entry p of tile
// t (the caller checks that no window is wider than 65,536 columns).
int device_window_offsets(hipStream_t s, int ntiles, const int* tile_row, const int* ptr, const int* idx,
                          const int* win, int64_t nnz, DevBuf& off16);

// inst[t] = the instance of the first row of tile t (inst_of[tile_row[t]]; 0 when inst_of is null).
int device_tile_instances(hipStream_t s, int ntiles, const int* tile_row, const int* inst_of, DevBuf& inst);

// Descriptors of staged tiles, 8 ints each: first row, end row, first entry (ptr[first row]),
// end entry, window start, window end (win), instance (inst), 0.
int device_stage_descs(hipStream_t s, int ntiles, const int* tile_row, const int* ptr, const int* win,
                       const int* inst, DevBuf& desc);

// Per instance (rows off[q]..off[q+1]-1): the tiles holding its first and last row.
int device_instance_tiles(hipStream_t s, int count, int ntiles, const int* tile_row, const int* off, DevBuf& first,
                          DevBuf& last);

// tbase[t*S + k] = ptrs[k*stride + tile_row[t]] for t = 0..ntiles: each tile's first entry in
// each of the S pointer arrays.
int device_tile_bases(hipStream_t s, int ntiles, int S, const int* tile_row, const int* ptrs, int64_t stride,
                      DevBuf& tbase);

// Column split of a CSR (rows x cols) into S blocks of ceil(cols/S) columns: block-major
// entries with absolute row pointers bp[k*(rows+1) + r], column indices idx and perm[q] = the
// CSR position of split entry q; entries keep their CSR order within each block and row.
int device_split(hipStream_t s, int rows, int cols, int64_t nnz, const int* rp, const int* ci, int S, DevBuf& bp,
                 DevBuf& idx, DevBuf& perm);

// Tile-relative row offsets of a column split: toff[k*rows + r] = ptrs[k*stride + r] -
// ptrs[k*stride + tile_row[t]] for the tile t holding row r (clamped to 65535: only a row of
// a long-row tile, which is its tile's first row with offset 0, could exceed it otherwise).
int device_tile_offsets(hipStream_t s, int rows, int S, int ntiles, const int* tile_row, const int* ptrs,
                        int64_t stride, DevBuf& toff);

// Sliced ELL of a CSR in slices of kEllSlice rows: per-slice entry offsets (int32) and widths
// (the slice's longest row); entry q of row r at off[s] + kEllSlice*q + (r mod kEllSlice);
// perm = CSR position or -1 for padding, whose index repeats the row's first column (0 for an
// empty row). *padding = ELL entries / nonzeros. The copy is built only when it fits 32-bit
// offsets and (forced or padding <= max_padding); *ell_len = its entries, else 0.
int device_ell(hipStream_t s, int rows, int64_t nnz, const int* rp, const int* ci, double max_padding, bool forced,
               double* padding, int64_t* ell_len, DevBuf& off, DevBuf& width, DevBuf& idx, DevBuf& perm);

}  // namespace vj
