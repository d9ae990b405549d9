/* vajra_pdhg.h -- C ABI of the MI355X (gfx950) restarted-PDHG LP engine.
 *
 * Plain C: POD structs, plain pointers and sizes, int status codes, no exceptions and no
 * torch types. The C++17 solver API (include/indus/, libindus_hip.so) calls only
 * these functions; any other host (ctypes, cgo, JNI) can bind them the same way.
 *
 * The reference has no C ABI: its device seam is the C++ function
 * indus::gpu::solve_pdhg_cuda_internal (src/solvers/gpu/pdhg_cuda.cu:109), reached via
 * indus::gpu::solve_pdhg_gpu (src/solvers/gpu/gpu_context.cpp:65). Each entry point below
 * names the part of that seam it replaces.
 *
 * Matrix layout conventions: the caller passes A as CSR of A^T (which is the CSC of A, the
 * reference Model layout, include/indus/sparse.hpp:20-65) and, optionally, also as CSR; when
 * the CSR arrays are NULL the engine builds the CSR of A on the device from the CSC (a
 * stable sort of the entries by row, the order SparseMatrixCSC::to_csr produces,
 * src/linalg/sparse.cpp:147-184). Row pointers are int64 as in the reference; the engine
 * narrows them to int32 on the device (nnz < 2^31 is checked). Indices must be strictly
 * ascending within each row and column.
 * Infinite bounds: any |bound| >= 1e20 (or +-inf) is infinite.
 *
 * Threading: one context is used by one host thread at a time. Calls on different
 * contexts may run concurrently. All work of a context goes to its own HIP stream.
 */
#ifndef VAJRA_PDHG_H
#define VAJRA_PDHG_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------------ */
enum {
    VJ_OK = 0,
    VJ_ERR_INVALID_ARGUMENT = 1,
    VJ_ERR_NO_DEVICE = 2,
    VJ_ERR_ALLOC = 3,
    VJ_ERR_LAUNCH = 4,       /* a HIP call or kernel launch failed */
    VJ_ERR_NOT_READY = 5,    /* e.g. solve before upload */
    VJ_ERR_UNSUPPORTED = 6
};
/* Human-readable text of the last error on this thread ("" if none). */
const char* vj_last_error(void);

/* ---- solver enums ------------------------------------------------------------ */
enum {
    /* Reference iteration map (src/solvers/pdhg/pdhg.cpp:196-366): vanilla
     * Chambolle-Pock, dual step first, ascending-index sums without fused multiply-add,
     * check every check_frequency iterations with the reference's stopping rule,
     * step rebalancing, stagnation restart and best iterate. Bit-reproduces the
     * reference CPU PDHG. Setup (Ruiz, power iteration) runs on the host as in
     * pdhg_cuda.cu:149-162. */
    VJ_MODE_REFERENCE = 0,
    /* Restarted reflected Halpern PDHG with adaptive restarts on the fixed-point
     * residual, primal-weight smoothing and device-side Ruiz + Pock-Chambolle +
     * bound/objective rescaling; stops on the relative KKT error (FP64). */
    VJ_MODE_RESTARTED = 1
};
enum {
    VJ_PREC_FP64 = 0,       /* FP64 everywhere */
    VJ_PREC_FP32 = 1,       /* FP32 iterates, matrix and projections; FP64 refinement */
    VJ_PREC_BF16 = 2,       /* BF16 matrix values, FP32 accumulate/iterates; FP64 refinement */
    VJ_PREC_FP16 = 3        /* FP16 matrix values (the scaled matrix must fit |a| <= 65504),
                               FP32 accumulate/iterates; FP64 refinement */
};
/* SolveStatus numbering of the reference (include/indus/types.hpp:34-47). */
enum {
    VJ_STATUS_NOT_SOLVED = 0, VJ_STATUS_OPTIMAL = 1, VJ_STATUS_INFEASIBLE = 2,
    VJ_STATUS_UNBOUNDED = 3, VJ_STATUS_FEASIBLE = 5, VJ_STATUS_ITERATION_LIMIT = 6,
    VJ_STATUS_TIME_LIMIT = 7, VJ_STATUS_NUMERICAL_ERROR = 9
};

/* ---- device probe (replaces probe_device, gpu_context.cpp:16-59) ----------- */
typedef struct vj_device_info {
    int available;
    int device_count;
    char name[256];
    char arch[64];                 /* e.g. "gfx950" */
    size_t total_bytes, free_bytes;
    int compute_units;
    int wavefront_size;
    int major, minor;
} vj_device_info;
int vj_device_probe(vj_device_info* out);

/* ---- parameters and results (replace PdhgOptions / DeviceDiagnostics /
 *      GpuTimingDiagnostics, include/indus/pdhg.hpp:17-43, gpu.hpp:27-46) --------- */
typedef struct vj_params {
    int mode;                      /* VJ_MODE_* */
    int precision;                 /* VJ_PREC_* (VJ_MODE_RESTARTED only) */
    int64_t max_iterations;        /* reference default 1e6 (options.hpp:26) */
    double time_limit;             /* seconds; checked at each check */
    double tolerance;              /* REFERENCE: abs prim/dual tol; RESTARTED: rel-KKT target */
    double gap_tolerance;          /* REFERENCE only */
    int check_frequency;           /* REFERENCE: iterations between checks (100);
                                      RESTARTED: iterations per check block, 0 = autotune */
    int restart_stagnation_limit;  /* REFERENCE: 200 */
    double step_size_safety;       /* REFERENCE: 0.95, clamped to [0.5, 0.99] */
    int enable_scaling;            /* Ruiz on/off */
    int ruiz_iterations;           /* RESTARTED: default 10 */
    int pock_chambolle;            /* RESTARTED: 1 = apply alpha=1 diagonal scaling */
    int bound_objective_rescaling; /* RESTARTED: 1 = scale b and c to unit-ish norm */
    double reflection;             /* RESTARTED: reflection rho in ((1+rho)T - rho I), [0,1] */
    double step_scale;             /* RESTARTED: eta = step_scale / ||A||_2 */
    int use_graph;                 /* RESTARTED: replay the check block as a hipGraph */
    int lanes_A;                   /* SpMV kernel for rows of A: 0 = autotune; 2..32 = lanes per row
                                      (lane groups, or lanes per block segment when split);
                                      -1 = tile-streamed; -2 = tile-streamed reading interleaved
                                      (index, value) records (FP32; other precisions as -1);
                                      -3 = sliced ELL (slices of 64 rows, single LP; where no
                                      ELL copy exists, as -1); -4 = sliced ELL reading (index,
                                      value) records (FP32; other precisions as -3; no ELL copy:
                                      as -2); -5 = LDS-staged (tiles of one instance copy the
                                      window of the gathered vector they read into LDS; built
                                      for batches, and for a single LP when forced); -6 = the
                                      same reading (index, value) records (FP32; otherwise as
                                      -5); -7 = LDS-staged reading 16-bit column offsets from
                                      each tile's window start (built where every window
                                      spans at most 65,536 columns; otherwise as -5). The
                                      result reports the kernel that ran with the
                                      same codes; the autotuner tries ELL where its padding is
                                      at most 25% of the nonzeros, and the staged kernels for
                                      batches */
    int lanes_At;                  /* same for rows of A^T */
    int refine_rounds;             /* FP32/BF16: max FP64 refinement rounds (0 = 64) */
    int verbose;
    int col_splits_A;              /* column blocks of A for the XCD-split SpMV: 0 = by vector
                                      size vs L2, 1 = fused single pass, 2/4/8 = forced */
    int col_splits_At;             /* same for A^T (blocks of rows of A) */
    /* RESTARTED: restart rule on the fixed-point residual r (r0 at the epoch start):
     * restart if r <= sufficient*r0, or r <= necessary*r0 and r grew since the last check,
     * or the epoch is longer than artificial * (total iterations). */
    double restart_sufficient;     /* default 0.2 */
    double restart_necessary;      /* default 0.8 */
    double restart_artificial;     /* default 0.5 (PDLP: 0.36) */
    double primal_weight_smoothing;/* proportional gain theta of the primal-weight update at a restart:
                                      with e = log(w |dx| / |dy|) (|dx|, |dy|: the distances the epoch
                                      moved), log w <- log w - theta e - ki (sum of e over the solve's
                                      restarts); with ki = 0 this is PDLP's smoothing
                                      log w <- theta log(|dy|/|dx|) + (1-theta) log w; default 0.5 */
    int power_iterations;          /* RESTARTED: power-iteration steps for ||A||_2; default 64 */
    double restart_drift;          /* RESTARTED: also restart (an FP64 refinement step) when a check's
                                      FP64 rel-KKT exceeds restart_drift times the least rel-KKT of
                                      the epoch; 0 = default (2 with FP32 values, 4 with BF16 and
                                      FP16 values refined at every check, off otherwise), < 0 = off,
                                      > 1 = this factor for every precision */
    double primal_weight_integral; /* RESTARTED: integral gain ki of the primal-weight update (see
                                      primal_weight_smoothing), in [0, 1); default 0 (PDLP's
                                      smoothing) */
    int refine_every_check;        /* RESTARTED: where the FP64 centre moves (the refinement step of
                                      FP32/BF16/FP16 iterates): 1 = at every check, keeping the
                                      Halpern epoch (its anchor becomes an explicit vector); -1 = at
                                      restarts only; 0 = default (every check with BF16 and FP16
                                      values, restarts only with FP32 and FP64) */
    int stage_threads;             /* RESTARTED: workgroup size of the LDS-staged kernels (lanes -5 /
                                      -6 / -7): 0 = the autotuner times 1,024 and 256 threads (a
                                      forced staged kernel runs 1,024), 256 or 1024 = that size */
} vj_params;
/* Fill defaults for a mode (reference defaults for VJ_MODE_REFERENCE). */
void vj_params_default(vj_params* p, int mode);

typedef struct vj_result {
    int status;                    /* VJ_STATUS_* */
    int converged;
    int64_t iterations;            /* reference quirk kept: max_iterations+1 at the limit */
    int restart_count;
    /* reference criterion at the last check (scaled space, inf-norm) -- REFERENCE */
    double primal_residual, dual_residual, duality_gap;
    /* FP64 audit of the returned point on the original data (build-defined rel-KKT) */
    double rel_kkt, rel_primal, rel_dual, rel_gap;
    double primal_objective, dual_objective;
    double spectral_norm, tau, sigma, primal_weight;
    double setup_time, iteration_time, total_time, upload_time, download_time;
    double time_to_target;         /* iteration seconds until rel-KKT <= tolerance, -1 if never */
    int64_t host_device_transfers;
    int refinements;               /* FP64 refinement rounds used (mixed precision) */
    int lanes_A, lanes_At, check_interval;
    int64_t checks;
    int col_splits_A, col_splits_At;   /* column blocks used by the XCD-split SpMV (1 = fused) */
    double layout_time;            /* seconds of this solve spent building kernel layouts on the
                                      device (column split, sliced ELL, their row tiles); reused
                                      by later solves of the same upload, then near 0 */
} vj_result;

/* ---- single-LP context (replaces solve_pdhg_cuda_internal, pdhg_cuda.cu:109-375) */
typedef struct vj_ctx vj_ctx;
int vj_ctx_create(int device, vj_ctx** out);
int vj_ctx_destroy(vj_ctx* ctx);

/* Host LP in the Model layout, optionally with the CSR of A (csr_* all NULL: built on the
 * device). Arrays are copied; the engine owns its device buffers (replaces the host CSR build
 * and the H2D uploads of pdhg_cuda.cu:159,187-231). */
typedef struct vj_lp {
    int m, n;
    int64_t nnz;
    const int64_t* csr_row_ptr;    /* m+1, or NULL (with csr_col_idx and csr_vals) */
    const int32_t* csr_col_idx;    /* nnz, ascending in each row */
    const double* csr_vals;        /* nnz */
    const int64_t* csc_col_ptr;    /* n+1 (= CSR of A^T) */
    const int32_t* csc_row_idx;    /* nnz, ascending in each column */
    const double* csc_vals;        /* nnz */
    const double* c;               /* n, objective as given */
    const double* col_lower;       /* n */
    const double* col_upper;       /* n */
    const double* row_lower;       /* m */
    const double* row_upper;       /* m */
    int maximize;                  /* 1: maximise c^T x (negated internally, pdhg.cpp:133-135) */
    double objective_offset;
} vj_lp;
int vj_problem_upload(vj_ctx* ctx, const vj_lp* lp);
int vj_params_set(vj_ctx* ctx, const vj_params* p);
/* Run the solver on the uploaded problem to termination. */
int vj_solve(vj_ctx* ctx, vj_result* out);
/* Download the returned point in the reference's output convention
 * (pdhg.cpp:380-418): x unscaled; row_dual = -y (minimise) / +y (maximise) of the PDHG
 * dual; any pointer may be NULL. y_canonical (optional) is the dual of the canonical
 * minimisation problem (y >= 0 prices lower row bounds), the one rel-KKT audits. */
int vj_download(vj_ctx* ctx, double* x, double* row_dual, double* y_canonical);
/* Reference-map trajectory probe: run exactly k iterations from the start (no checks if
 * check_frequency exceeds k) and download the current scaled iterates. REFERENCE only. */
int vj_reference_iterates(vj_ctx* ctx, int64_t k, double* x, double* x_bar, double* y);

/* FP64 relative KKT of an arbitrary point (x, canonical y) on the uploaded problem,
 * computed on the device. out: rel_kkt, rel_p, rel_d, rel_gap, pobj, dobj, |r_p|, |r_d|. */
int vj_audit_point(vj_ctx* ctx, const double* x, const double* y_canonical, double* out8);

/* The cumulative equilibration factors of the last VJ_MODE_RESTARTED solve, computed on the
 * device: row_scale[m] and col_scale[n] after the Ruiz passes and the Pock-Chambolle pass
 * (before the per-instance bound/objective rescaling); A_scaled = diag(row) A diag(col).
 * Either pointer may be NULL. (Counterpart of RuizScaler's factors, src/linalg/scaling.hpp:11-17,
 * which the reference computes on the host, src/linalg/scaling.cpp:7-153.) */
int vj_scaling_factors(vj_ctx* ctx, double* row_scale, double* col_scale);

/* Timing of the dominant kernels of the last solve, measured with HIP events on the
 * context stream: kernel names, average microseconds per launch, launches, and
 * algorithmic bytes per launch. Returns the number of entries written (<= cap). */
typedef struct vj_kernel_stat {
    char name[48];
    double avg_us;
    int64_t launches;
    double bytes_per_launch;
} vj_kernel_stat;
int vj_kernel_stats(vj_ctx* ctx, vj_kernel_stat* out, int cap);
/* Time `reps` back-to-back launches of each fused iteration kernel on the current
 * iterates (after a solve; iterates are restored). Fills stats as vj_kernel_stats. */
int vj_benchmark_kernels(vj_ctx* ctx, int reps, vj_kernel_stat* out, int cap);
/* Time check blocks of the last VJ_MODE_RESTARTED solve on its device-resident problem: one
 * block is check_interval iterations, the FP64 check, the control kernel and the conditional
 * restart, replayed from the solve's hipGraph. Every instance iterates through all blocks
 * (none may stop in between); the final state of the solve is restored afterwards and the
 * downloaded point is that of the solve. `warmup` untimed blocks, then `steps` blocks timed
 * with HIP events on the context stream: ms_per_block = their average, iterations_per_block
 * = the check interval. (Benchmark entry; the reference times whole solves only,
 * GpuTimingDiagnostics, include/indus/gpu.hpp:37-46.) */
int vj_benchmark_blocks(vj_ctx* ctx, int warmup, int steps, double* ms_per_block, int* iterations_per_block);

/* ---- batched LPs (new functionality; no reference counterpart) --------------- */
/* Instances are packed block-diagonally into one CSR and one CSR(A^T); each instance
 * keeps its own step sizes, primal weight, restart and convergence state. Arrays are
 * concatenations of the per-instance arrays with per-instance index bases: row/col
 * indices are LOCAL to the instance, pointers are local offsets (each instance's
 * pointer array starts at 0). The CSR arrays may be NULL together, as for one LP. */
typedef struct vj_batch {
    int count;
    const int32_t* m;              /* count */
    const int32_t* n;              /* count */
    const int64_t* nnz;            /* count */
    const int64_t* csr_row_ptr;    /* sum(m+1) */
    const int32_t* csr_col_idx;    /* sum(nnz) */
    const double* csr_vals;
    const int64_t* csc_col_ptr;    /* sum(n+1) */
    const int32_t* csc_row_idx;
    const double* csc_vals;
    const double* c;               /* sum(n), objective as given (see maximize) */
    const double* col_lower;
    const double* col_upper;
    const double* row_lower;       /* sum(m) */
    const double* row_upper;
    const int32_t* maximize;       /* count, 1: instance maximises c^T x (negated internally, as in
                                      vj_lp); NULL: every instance minimises */
    const double* objective_offset;/* count, added to the instance's objective (it enters the
                                      relative gap of the stopping test); NULL: all zero */
} vj_batch;
int vj_batch_upload(vj_ctx* ctx, const vj_batch* b);
/* Solve every instance; per-instance results go to out[count]. */
int vj_batch_solve(vj_ctx* ctx, vj_result* out);
/* x: sum(n), y_canonical: sum(m); either may be NULL. */
int vj_batch_download(vj_ctx* ctx, double* x, double* y_canonical);

/* ---- kernel-level operators on caller device memory ----------------------------
 * Device pointers, a hipStream_t passed as void* (NULL = default stream). These restate
 * the reference's separate CUDA kernels (src/solvers/gpu/spmv_kernels.cuh) one by one, with
 * their contracts, so that each can be checked on its own against the reference's kernel
 * tests. They are not the restarted engine's hot loop: vj_solve runs fused kernels (SpMV,
 * projection, extrapolation and Halpern step in one launch per half-step, over tiled,
 * column-split or sliced-ELL layouts), which these entry points do not expose. */
/* y = A x over CSR (int64 row pointers as in the reference); lanes = lanes per row.
 * Replaces spmv_csr_warp_kernel (spmv_kernels.cuh:39-76) and, applied to the CSR of A^T,
 * spmv_csc_transpose_warp_kernel (:101-138). exact_order = 1: one lane per row, ascending
 * sum, no FMA, skip |x_j| <= 1e-11 when skip_tiny = 1 (sparse.cpp:207). */
int vj_op_spmv(int rows, const int64_t* row_ptr, const int32_t* col_idx, const double* vals,
               const double* x, double* y, int lanes, int exact_order, int skip_tiny,
               void* stream);
/* Reference dual step y = yhat - sigma*clamp(yhat/sigma, l, u), yhat = y + sigma*w
 * (pdhg.cpp:203-214; replaces dual_update_kernel, spmv_kernels.cuh:163-180). */
int vj_op_dual_update(int m, double sigma, const double* w, const double* row_lower,
                      const double* row_upper, double* y, void* stream);
/* Reference primal step x' = clamp(x - tau (v + c), l, u), x_bar = 2x' - x
 * (pdhg.cpp:228-241; replaces primal_update_kernel, spmv_kernels.cuh:185-206). */
int vj_op_primal_update(int n, double tau, const double* v, const double* c,
                        const double* col_lower, const double* col_upper, double* x,
                        double* x_bar, void* stream);
/* Deterministic diagnostics (replaces compute_diagnostics_device_kernel,
 * spmv_kernels.cuh:210-278): out[0] max row/col primal violation given Ax,
 * out[1] max dual violation (reference thresholds 1e-5/1e-6, pdhg.cpp:298-306),
 * out[2] objective c^T x, out[3] NaN/Inf flag. out is a host pointer; synchronous. */
int vj_op_diagnostics(int m, int n, const double* Ax, const double* row_lower,
                      const double* row_upper, const double* Aty, const double* c,
                      const double* x, const double* col_lower, const double* col_upper,
                      double* out4, void* stream);
/* vj_op_diagnostics without any allocation: no counterpart in the reference. For callers
 * that cannot allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc). Same inputs and the same four results, bit for bit, but
 * out4_device is a DEVICE pointer and the call is asynchronous on `stream` (nothing is
 * copied to the host, nothing is waited for). The per-workgroup partials live in a static
 * device array of 1 MiB per GPU: max(m, n) must be <= 8,388,608, else
 * VJ_ERR_INVALID_ARGUMENT. Every call on one GPU shares that array, so calls that overlap
 * on one GPU (different streams or host threads) race; calls ordered on one stream do not. */
int vj_op_diagnostics_static(int m, int n, const double* Ax, const double* row_lower,
                             const double* row_upper, const double* Aty, const double* c,
                             const double* x, const double* col_lower, const double* col_upper,
                             double* out4_device, void* stream);

#ifdef __cplusplus
}
#endif
#endif
