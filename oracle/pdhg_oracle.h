/* TEST INFRASTRUCTURE -- NOT PART OF THE PRODUCT.
 *
 * CPU restatement of the reference's restarted-PDHG LP path, used only for comparison
 * by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg. The product
 * (libvajra_hip.so / libindus_hip.so) never links, loads or calls this code.
 *
 * Every function cites the reference file:line it restates. Arithmetic follows the
 * reference's operation order exactly and the library is compiled with
 * -ffp-contract=off, so results are bit-identical to the reference built without FMA.
 *
 * Pinning (what checks each function; see tests/test_oracle.py):
 *   or_csc_multiply, or_csc_multiply_transpose : reference KATs (test_cuda_kernels.cu:74-140,
 *       test_pdhg.cpp:442-477), dense brute force on random inputs, oracle/_ref bit-exact.
 *   or_csc_to_csr : dense round trip, oracle/_ref-compatible ordering (ascending columns).
 *   or_ruiz : row/col inf-norm invariant (test_linalg.cpp:283-311) + oracle/_ref bit-exact.
 *   or_spectral_norm : closed forms (diagonal, rank one), numpy SVD bound, oracle/_ref bit-exact.
 *   or_pdhg_solve : reference LP KATs (test_pdhg.cpp:54-314), golden trajectories written
 *       by oracle/_ref (tests/golden), bit-exact.
 *   or_recompute_quality, or_verify_solution : constructed points with known violations,
 *       oracle/_ref on the same points.
 *   or_rel_kkt : build-defined metric (no reference counterpart); pinned by a known optimal
 *       primal/dual pair (0), hand-computed residuals on tiny LPs, and invariance checks.
 */
#ifndef VAJRA_PDHG_ORACLE_H
#define VAJRA_PDHG_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* LP in the reference's Model layout (include/indus/model.hpp:16-55): A in CSC. */
typedef struct {
    int m, n;
    const int64_t* col_ptr;  /* n+1 */
    const int32_t* row_idx;  /* nnz, ascending within a column */
    const double* vals;      /* nnz */
    const double* c;         /* n */
    const double* col_lower; /* n */
    const double* col_upper; /* n */
    const double* row_lower; /* m */
    const double* row_upper; /* m */
    int maximize;
    double offset;
} or_lp;

/* PdhgOptions (include/indus/pdhg.hpp:17-28) plus test-only controls. */
typedef struct {
    int64_t max_iterations;
    double time_limit;
    double primal_tolerance;
    double dual_tolerance;
    double gap_tolerance;
    int check_frequency;
    int restart_stagnation_limit;
    double step_size_safety;
    int enable_scaling;
    /* test-only: shadow rel-KKT monitor target at each check (0 = off); it reads the
     * current iterate and never changes the trajectory. */
    double kkt_target;
    /* test-only: >0 stops after this iteration, before finalisation, and returns the
     * current scaled iterates in traj_*. */
    int64_t stop_at_iter;
} or_pdhg_opts;

typedef struct {
    int status; /* SolveStatus numbering of include/indus/types.hpp:34-47 */
    int converged;
    int64_t iterations;
    int restart_count;
    double primal_residual, dual_residual, duality_gap;
    double spectral_norm, tau, sigma;
    double setup_time, iteration_time, total_time;
    double objective;
    /* shadow monitor */
    double kkt_first_time;   /* seconds of PDHG work (monitor time excluded); -1 if never */
    int64_t kkt_first_iter;  /* -1 if never */
    double kkt_last;         /* rel-KKT at the last check */
    double monitor_time;
    /* step sizes at the end (tau/sigma above are the initial ones, as PdhgDiagnostics
     * records them at pdhg.cpp:159-160) */
    double final_tau, final_sigma;
} or_pdhg_result;

void or_default_opts(or_pdhg_opts* o);

/* sparse.cpp:196-217 -- y = A x by column scatter, skipping |x_j| <= 1e-11. */
void or_csc_multiply(int m, int n, const int64_t* col_ptr, const int32_t* row_idx,
                     const double* vals, const double* x, double* y);
/* sparse.cpp:219-238 -- v_j = sum_i a_ij y_i, ascending i. */
void or_csc_multiply_transpose(int m, int n, const int64_t* col_ptr, const int32_t* row_idx,
                               const double* vals, const double* y, double* v);
/* sparse.cpp:147-184 -- counting-sort CSC -> CSR (columns ascending within a row). */
void or_csc_to_csr(int m, int n, const int64_t* col_ptr, const int32_t* row_idx,
                   const double* vals, int64_t* row_ptr, int32_t* col_idx, double* csr_vals);
/* scaling.cpp:7-153 -- Ruiz l_inf equilibration, in place. Returns 1 if scaled. */
int or_ruiz(int m, int n, const int64_t* col_ptr, const int32_t* row_idx, double* vals,
            double* c, double* row_lower, double* row_upper, double* col_lower,
            double* col_upper, int max_iters, double tol, double* row_scale, double* col_scale);
/* pdhg.cpp:11-56 -- power iteration on A^T A. */
double or_spectral_norm(int m, int n, const int64_t* col_ptr, const int32_t* row_idx,
                        const double* vals, int max_iter, double tol);
/* pdhg.cpp:70-453 -- CpuPdhgSolver::solve. Outputs follow Solution (model.hpp:71-116). */
int or_pdhg_solve(const or_lp* lp, const or_pdhg_opts* opts, double* x_out,
                  double* row_dual_out, double* col_dual_out, double* row_value_out,
                  or_pdhg_result* res, double* traj_x, double* traj_xbar, double* traj_y);

/* model.cpp:203-374 (LP branch) -- Solution::recompute_quality.
 * q[0] max_primal_violation, q[1] max_dual_violation, q[2] max_complementarity_violation,
 * q[3] is_primal_feasible, q[4] is_dual_feasible. col_dual_out may be NULL. */
void or_recompute_quality(const or_lp* lp, const double* x, const double* row_dual,
                          double* col_dual_out, double* q);
/* verifier.cpp:25-330 (LP branch) -- verify_solution.
 * v[0] passed, 1 max_primal_violation, 2 max_dual_violation, 3 max_complementarity_violation,
 * 4 objective_discrepancy, 5 recomputed_objective, 6 primal_feasible, 7 dual_feasible,
 * 8 bounds_feasible, 9 objective_matches. row_dual may be NULL. */
void or_verify_solution(const or_lp* lp, const double* x, const double* row_dual,
                        double reported_objective, double tol, double* v);

/* Build-defined relative KKT error of (x, y) for the canonical minimisation problem
 * (c' = sense*c, offset' = sense*offset; y >= 0 prices lower row bounds), evaluated in FP64
 * on the original data. out[0] rel_kkt, 1 rel_primal, 2 rel_dual, 3 rel_gap, 4 primal_obj, 5 dual_obj,
 * 6 |r_p|_2, 7 |r_d|_2. */
void or_rel_kkt(const or_lp* lp, const double* x, const double* y, double* out);

#ifdef __cplusplus
}
#endif
#endif
