/* Seeded synthetic LP generators for the PDHG workloads (random equality LPs and
 * block-staircase multi-period LPs). Host-only code in libvajra_hip.so; no GPU needed.
 *
 * Counterpart in the reference: apps/benchmark_runner.cpp:113-157
 * (generate_synthetic_sparse_lp, an LCG generator). The distributions here differ on
 * purpose: they produce LPs that are feasible (b = A x0 with x0 inside the box) and
 * bounded, so every instance has an optimum to converge to.
 *
 * Random streams: SplitMix64 seeds xoshiro256** (Vigna, public domain);
 * U[0,1) = (u64 >> 11) * 2^-53; N(0,1) by Box-Muller; k distinct rows per column by
 * Floyd's sampling, stored ascending. All outputs use the reference's Model layout:
 * A in CSC (col_ptr int64[n+1], row_idx int32[nnz], vals f64[nnz]). Infinite bounds
 * are +-1e20, as the reference's tests write them.
 */
#ifndef VAJRA_LPGEN_H
#define VAJRA_LPGEN_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Random LP: m equality rows, n columns, k nnz per column at distinct uniform rows,
 * a_ij ~ N(0,1). A fraction `scaled_frac` of the columns (chosen by Bernoulli draws)
 * is multiplied by 10^U[-log10_range, log10_range] before b and c are formed.
 * x0 ~ U[0,1]^n, b = A x0, 0 <= x <= 10, y0 ~ N(0,1)^m, s0 ~ U[0,1]^n, c = A^T y0 + s0.
 * nnz = n*k. Returns 0 on success, -1 on bad arguments. x0/y0 may be NULL. */
int vj_gen_random_lp(uint64_t seed, int m, int n, int k, double scaled_frac,
                     double log10_range, int64_t* col_ptr, int32_t* row_idx, double* vals,
                     double* c, double* col_lower, double* col_upper, double* row_lower,
                     double* row_upper, double* x0, double* y0);

/* Block-staircase LP: T periods, each a rows_pp x cols_pp block with k nnz per column
 * (N(0,1)); the last `inv` columns of period t also carry a +1 entry in row j' of
 * period t+1 (identity inventory coupling, j' = inventory column index 0..inv-1).
 * c ~ U[1,10]; 0 <= x <= U[5,50]; x0 ~ U[0,5]; b = A x0 (equality rows).
 * m = T*rows_pp, n = T*cols_pp, nnz = n*k + inv*(T-1). */
int vj_gen_staircase_lp(uint64_t seed, int periods, int rows_pp, int cols_pp, int k, int inv,
                        int64_t* col_ptr, int32_t* row_idx, double* vals, double* c,
                        double* col_lower, double* col_upper, double* row_lower,
                        double* row_upper, double* x0);

/* `count` integers uniform on [lo, hi] from the stream seeded by `seed`. */
int vj_gen_uniform_ints(uint64_t seed, int count, int lo, int hi, int32_t* out);

#ifdef __cplusplus
}
#endif
#endif
