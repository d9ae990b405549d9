/* TEST INFRASTRUCTURE -- NOT PART OF THE PRODUCT. See pdhg_oracle.h for what this is,
 * who may call it, and how each function is pinned.
 *
 * Operation order follows the reference statement by statement; build with
 * -ffp-contract=off (oracle/Makefile) so no multiply-add is fused. */
#define _POSIX_C_SOURCE 199309L
#include "pdhg_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

/* std::max / std::min semantics: max(a,b) = (a < b) ? b : a. */
#define SMAX(a, b) (((a) < (b)) ? (b) : (a))
#define SMIN(a, b) (((b) < (a)) ? (b) : (a))

static const double kZeroDrop = 1e-11;         /* tolerances.hpp:28 */
static const double kPrimalFeas = 1e-7;        /* tolerances.hpp:7 */
static const double kDualFeas = 1e-7;          /* tolerances.hpp:10 */

enum { ST_NOT_SOLVED = 0, ST_OPTIMAL = 1, ST_UNBOUNDED = 3, ST_FEASIBLE = 5,
       ST_ITERATION_LIMIT = 6, ST_TIME_LIMIT = 7, ST_NUMERICAL_ERROR = 9 };

static double now_sec(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

static double* dalloc(size_t k) { return (double*)calloc(k ? k : 1, sizeof(double)); }

void or_default_opts(or_pdhg_opts* o) {
    /* pdhg.hpp:17-28 defaults, with the ctor mapping of pdhg.cpp:58-68 applied to a
     * default Options (options.hpp:25-37): max_iterations <- iteration_limit = 1e6. */
    o->max_iterations = 1000000;
    o->time_limit = 1e20;
    o->primal_tolerance = 1e-6;
    o->dual_tolerance = 1e-6;
    o->gap_tolerance = 1e-6;
    o->check_frequency = 100;
    o->restart_stagnation_limit = 200;
    o->step_size_safety = 0.95;
    o->enable_scaling = 1;
    o->kkt_target = 0.0;
    o->stop_at_iter = 0;
}

/* sparse.cpp:196-217 with alpha = 1, beta = 0. */
void or_csc_multiply(int m, int n, const int64_t* col_ptr, const int32_t* row_idx,
                     const double* vals, const double* x, double* y) {
    for (int i = 0; i < m; ++i) y[i] = 0.0;
    for (int j = 0; j < n; ++j) {
        const double xj = x[j];
        if (fabs(xj) <= kZeroDrop) continue;
        const double alpha_xj = 1.0 * xj;
        for (int64_t k = col_ptr[j]; k < col_ptr[j + 1]; ++k) {
            const double prod = alpha_xj * vals[k];
            y[row_idx[k]] = y[row_idx[k]] + prod;
        }
    }
}

/* sparse.cpp:219-238 with alpha = 1, beta = 0. */
void or_csc_multiply_transpose(int m, int n, const int64_t* col_ptr, const int32_t* row_idx,
                               const double* vals, const double* y, double* v) {
    (void)m;
    for (int j = 0; j < n; ++j) {
        double dot = 0.0;
        for (int64_t k = col_ptr[j]; k < col_ptr[j + 1]; ++k) {
            const double prod = vals[k] * y[row_idx[k]];
            dot = dot + prod;
        }
        const double scaled = 1.0 * dot;
        v[j] = 0.0 + scaled;
    }
}

/* sparse.cpp:147-184. */
void or_csc_to_csr(int m, int n, const int64_t* col_ptr, const int32_t* row_idx,
                   const double* vals, int64_t* row_ptr, int32_t* col_idx, double* csr_vals) {
    int64_t* cursor = (int64_t*)calloc((size_t)m + 1, sizeof(int64_t));
    for (int i = 0; i <= m; ++i) row_ptr[i] = 0;
    const int64_t nnz = col_ptr[n];
    for (int64_t k = 0; k < nnz; ++k) row_ptr[row_idx[k] + 1]++;
    for (int i = 0; i < m; ++i) row_ptr[i + 1] += row_ptr[i];
    memcpy(cursor, row_ptr, sizeof(int64_t) * ((size_t)m + 1));
    for (int j = 0; j < n; ++j)
        for (int64_t k = col_ptr[j]; k < col_ptr[j + 1]; ++k) {
            const int64_t dest = cursor[row_idx[k]]++;
            col_idx[dest] = j;
            csr_vals[dest] = vals[k];
        }
    free(cursor);
}

/* scaling.cpp:7-153. */
int or_ruiz(int m, int n, const int64_t* col_ptr, const int32_t* row_idx, double* vals,
            double* c, double* row_lower, double* row_upper, double* col_lower,
            double* col_upper, int max_iters, double tol, double* row_scale, double* col_scale) {
    if (m <= 0 || n <= 0) return 0;
    double* r_k = dalloc((size_t)m);
    double* c_k = dalloc((size_t)n);
    double* row_inf = dalloc((size_t)m);
    double* col_inf = dalloc((size_t)n);
    for (int i = 0; i < m; ++i) { row_scale[i] = 1.0; r_k[i] = 1.0; }
    for (int j = 0; j < n; ++j) { col_scale[j] = 1.0; c_k[j] = 1.0; }
    for (int iter = 0; iter < max_iters; ++iter) {
        for (int i = 0; i < m; ++i) row_inf[i] = 0.0;
        for (int j = 0; j < n; ++j)
            for (int64_t k = col_ptr[j]; k < col_ptr[j + 1]; ++k) {
                const int r = row_idx[k];
                row_inf[r] = SMAX(row_inf[r], fabs(vals[k]));
            }
        for (int j = 0; j < n; ++j) {
            col_inf[j] = 0.0;
            for (int64_t k = col_ptr[j]; k < col_ptr[j + 1]; ++k)
                col_inf[j] = SMAX(col_inf[j], fabs(vals[k]));
        }
        double max_row_err = 0.0, max_col_err = 0.0;
        for (int i = 0; i < m; ++i)
            if (row_inf[i] > kZeroDrop) max_row_err = SMAX(max_row_err, fabs(1.0 - row_inf[i]));
        for (int j = 0; j < n; ++j)
            if (col_inf[j] > kZeroDrop) max_col_err = SMAX(max_col_err, fabs(1.0 - col_inf[j]));
        if (max_row_err <= tol && max_col_err <= tol) break;
        for (int i = 0; i < m; ++i) {
            r_k[i] = (row_inf[i] > kZeroDrop) ? 1.0 / sqrt(row_inf[i]) : 1.0;
            row_scale[i] *= r_k[i];
        }
        for (int j = 0; j < n; ++j) {
            c_k[j] = (col_inf[j] > kZeroDrop) ? 1.0 / sqrt(col_inf[j]) : 1.0;
            col_scale[j] *= c_k[j];
        }
        for (int j = 0; j < n; ++j) {
            const double cs = c_k[j];
            for (int64_t k = col_ptr[j]; k < col_ptr[j + 1]; ++k) {
                const double f = r_k[row_idx[k]] * cs;
                vals[k] *= f;
            }
        }
    }
    /* scaling.cpp:103-149: c <- C c; finite row bounds x R; finite col bounds x C^-1. */
    for (int j = 0; j < n; ++j) c[j] *= col_scale[j];
    for (int i = 0; i < m; ++i) if (row_lower[i] > -1e20) row_lower[i] *= row_scale[i];
    for (int i = 0; i < m; ++i) if (row_upper[i] < 1e20) row_upper[i] *= row_scale[i];
    for (int j = 0; j < n; ++j) {
        const double inv = 1.0 / col_scale[j];
        if (col_lower[j] > -1e20) col_lower[j] *= inv;
    }
    for (int j = 0; j < n; ++j) {
        const double inv = 1.0 / col_scale[j];
        if (col_upper[j] < 1e20) col_upper[j] *= inv;
    }
    free(r_k); free(c_k); free(row_inf); free(col_inf);
    return 1;
}

/* pdhg.cpp:11-56. */
double or_spectral_norm(int m, int n, const int64_t* col_ptr, const int32_t* row_idx,
                        const double* vals, int max_iter, double tol) {
    if (m == 0 || n == 0 || col_ptr[n] == 0) return 1.0;
    double* v = dalloc((size_t)n);
    double* w = dalloc((size_t)m);
    double* u = dalloc((size_t)n);
    const double v0 = 1.0 / sqrt((double)n);
    for (int j = 0; j < n; ++j) v[j] = v0;
    double lambda_prev = 0.0, lambda = 1.0, result;
    int early = 0;
    for (int iter = 0; iter < max_iter; ++iter) {
        or_csc_multiply(m, n, col_ptr, row_idx, vals, v, w);
        or_csc_multiply_transpose(m, n, col_ptr, row_idx, vals, w, u);
        double norm_sq = 0.0;
        for (int j = 0; j < n; ++j) { const double p = u[j] * u[j]; norm_sq = norm_sq + p; }
        const double norm = sqrt(norm_sq);
        if (norm < 1e-15 || isnan(norm) || isinf(norm)) {
            result = (lambda > 1e-6) ? sqrt(lambda) : 1e-3;
            early = 1;
            break;
        }
        lambda = norm;
        if (iter > 0 && fabs(lambda - lambda_prev) <= tol * lambda) break;
        lambda_prev = lambda;
        const double inv_norm = 1.0 / norm;
        for (int j = 0; j < n; ++j) v[j] = u[j] * inv_norm;
    }
    if (!early) {
        const double sigma_max = sqrt(lambda);
        result = (sigma_max > 1e-12) ? sigma_max : 1e-6;
    }
    free(v); free(w); free(u);
    return result;
}

/* Build-defined rel-KKT (no reference counterpart). Infinite bounds are |b| >= 1e20. */
void or_rel_kkt(const or_lp* lp, const double* x, const double* y, double* out) {
    const int m = lp->m, n = lp->n;
    const double sense = lp->maximize ? -1.0 : 1.0;
    double* ax = dalloc((size_t)m);
    double* aty = dalloc((size_t)n);
    /* plain CSC products without the reference's skip rule: the metric audits the point */
    for (int j = 0; j < n; ++j)
        for (int64_t k = lp->col_ptr[j]; k < lp->col_ptr[j + 1]; ++k)
            ax[lp->row_idx[k]] += lp->vals[k] * x[j];
    for (int j = 0; j < n; ++j) {
        double dot = 0.0;
        for (int64_t k = lp->col_ptr[j]; k < lp->col_ptr[j + 1]; ++k)
            dot += lp->vals[k] * y[lp->row_idx[k]];
        aty[j] = dot;
    }
    double rp2 = 0.0, rd2 = 0.0, b2 = 0.0, c2 = 0.0, pobj = sense * lp->offset, dobj = sense * lp->offset;
    for (int i = 0; i < m; ++i) {
        const double l = lp->row_lower[i], u = lp->row_upper[i];
        const int lf = l > -1e20, uf = u < 1e20;
        double r = 0.0;
        if (lf && ax[i] < l) r = ax[i] - l;
        if (uf && ax[i] > u) r = ax[i] - u;
        rp2 += r * r;
        double bm = 0.0;
        if (lf) bm = fabs(l);
        if (uf && fabs(u) > bm) bm = fabs(u);
        b2 += bm * bm;
        /* row dual: y_i > 0 prices the lower bound, y_i < 0 the upper bound; a sign
         * with no finite bound to price is a dual residual */
        const double yi = y[i];
        if (yi > 0.0) { if (lf) dobj += yi * l; else rd2 += yi * yi; }
        else if (yi < 0.0) { if (uf) dobj += yi * u; else rd2 += yi * yi; }
    }
    for (int j = 0; j < n; ++j) {
        const double cj = sense * lp->c[j];
        c2 += cj * cj;
        pobj += cj * x[j];
        const double d = cj - aty[j];
        const double l = lp->col_lower[j], u = lp->col_upper[j];
        const int lf = l > -1e20, uf = u < 1e20;
        double lam = 0.0;
        if (d > 0.0 && lf) lam = d;
        else if (d < 0.0 && uf) lam = d;
        const double rd = d - lam;
        rd2 += rd * rd;
        if (lam > 0.0) dobj += lam * l;
        else if (lam < 0.0) dobj += lam * u;
    }
    const double rel_p = sqrt(rp2) / (1.0 + sqrt(b2));
    const double rel_d = sqrt(rd2) / (1.0 + sqrt(c2));
    const double rel_g = fabs(pobj - dobj) / (1.0 + fabs(pobj) + fabs(dobj));
    double k = rel_p;
    if (rel_d > k) k = rel_d;
    if (rel_g > k) k = rel_g;
    out[0] = k; out[1] = rel_p; out[2] = rel_d; out[3] = rel_g;
    out[4] = pobj; out[5] = dobj; out[6] = sqrt(rp2); out[7] = sqrt(rd2);
    free(ax); free(aty);
}

/* pdhg.cpp:70-440. */
int or_pdhg_solve(const or_lp* lp, const or_pdhg_opts* opts, double* x_out,
                  double* row_dual_out, double* col_dual_out, double* row_value_out,
                  or_pdhg_result* res, double* traj_x, double* traj_xbar, double* traj_y) {
    const double start_time = now_sec();
    memset(res, 0, sizeof(*res));
    res->kkt_first_time = -1.0;
    res->kkt_first_iter = -1;
    const int m = lp->m, n = lp->n;
    const int is_max = lp->maximize != 0;
    for (int j = 0; j < n; ++j) { x_out[j] = 0.0; if (col_dual_out) col_dual_out[j] = 0.0; }
    for (int i = 0; i < m; ++i) { row_dual_out[i] = 0.0; if (row_value_out) row_value_out[i] = 0.0; }

    /* pdhg.cpp:88-94 */
    if (n == 0) {
        res->status = ST_OPTIMAL;
        res->objective = lp->offset;
        res->converged = 1;
        return 0;
    }
    /* pdhg.cpp:96-127 */
    if (m == 0) {
        double obj = lp->offset;
        for (int j = 0; j < n; ++j) {
            const double cj = is_max ? -lp->c[j] : lp->c[j];
            const double lj = lp->col_lower[j], uj = lp->col_upper[j];
            if (cj > kDualFeas) {
                if (lj <= -1e20) { res->status = ST_UNBOUNDED; return 0; }
                x_out[j] = lj;
            } else if (cj < -kDualFeas) {
                if (uj >= 1e20) { res->status = ST_UNBOUNDED; return 0; }
                x_out[j] = uj;
            } else {
                x_out[j] = (lj > 0.0) ? lj : ((uj < 0.0) ? uj : 0.0);
            }
            obj += lp->c[j] * x_out[j];
        }
        res->status = ST_OPTIMAL;
        res->objective = obj;
        res->converged = 1;
        return 0;
    }

    /* pdhg.cpp:129-139 working copies */
    const int64_t nnz = lp->col_ptr[n];
    double* A = dalloc((size_t)nnz);
    memcpy(A, lp->vals, sizeof(double) * (size_t)nnz);
    double* wc = dalloc((size_t)n);
    double* wcl = dalloc((size_t)n);
    double* wcu = dalloc((size_t)n);
    double* wrl = dalloc((size_t)m);
    double* wru = dalloc((size_t)m);
    for (int j = 0; j < n; ++j) {
        wc[j] = is_max ? -lp->c[j] : lp->c[j];
        wcl[j] = lp->col_lower[j];
        wcu[j] = lp->col_upper[j];
    }
    for (int i = 0; i < m; ++i) { wrl[i] = lp->row_lower[i]; wru[i] = lp->row_upper[i]; }
    double* rs = dalloc((size_t)m);
    double* cs = dalloc((size_t)n);
    int scaled = 0;

    const double setup_start = now_sec();
    if (opts->enable_scaling)
        scaled = or_ruiz(m, n, lp->col_ptr, lp->row_idx, A, wc, wrl, wru, wcl, wcu, 10, 0.05, rs, cs);
    const double spectral_norm = or_spectral_norm(m, n, lp->col_ptr, lp->row_idx, A, 25, 1e-4);
    res->spectral_norm = spectral_norm;
    /* pdhg.cpp:156 std::clamp(safety, 0.5, 0.99) */
    double safety = opts->step_size_safety;
    safety = (safety < 0.5) ? 0.5 : ((0.99 < safety) ? 0.99 : safety);
    double tau = safety / spectral_norm;
    double sigma = safety / spectral_norm;
    res->tau = tau;
    res->sigma = sigma;

    double* x = dalloc((size_t)n);
    double* xb = dalloc((size_t)n);
    double* y = dalloc((size_t)m);
    for (int j = 0; j < n; ++j) {
        double val = 0.0;
        if (val < wcl[j]) val = wcl[j];
        if (val > wcu[j]) val = wcu[j];
        x[j] = val;
        xb[j] = val;
    }
    double* w = dalloc((size_t)m);
    double* v = dalloc((size_t)n);
    double* Ax = dalloc((size_t)m);
    double* best_x = dalloc((size_t)n);
    double* best_y = dalloc((size_t)m);
    memcpy(best_x, x, sizeof(double) * (size_t)n);
    memcpy(best_y, y, sizeof(double) * (size_t)m);
    double best_infeas = 1e20;
    int stagnation = 0;
    res->setup_time = now_sec() - setup_start;

    /* shadow monitor buffers (original space, canonical-min dual) */
    double* mx = opts->kkt_target > 0.0 ? dalloc((size_t)n) : NULL;
    double* my = opts->kkt_target > 0.0 ? dalloc((size_t)m) : NULL;

    const double iter_start = now_sec();
    int converged = 0;
    int status = ST_NOT_SOLVED;
    int64_t iter = 0;
    const int64_t max_iter = opts->max_iterations;
    const int check_freq = opts->check_frequency > 1 ? opts->check_frequency : 1;

    for (iter = 1; iter <= max_iter; ++iter) {
        /* pdhg.cpp:197-220 */
        or_csc_multiply(m, n, lp->col_ptr, lp->row_idx, A, xb, w);
        int has_nan = 0;
        for (int i = 0; i < m; ++i) {
            const double sw = sigma * w[i];
            const double y_hat = y[i] + sw;
            const double li = wrl[i], ui = wru[i];
            const double val = y_hat / sigma;
            const double clamped = (val < li) ? li : ((val > ui) ? ui : val);
            const double sc = sigma * clamped;
            const double y_new = y_hat - sc;
            if (isnan(y_new) || isinf(y_new)) has_nan = 1;
            y[i] = y_new;
        }
        if (has_nan) { status = ST_NUMERICAL_ERROR; break; }
        /* pdhg.cpp:222-247 */
        or_csc_multiply_transpose(m, n, lp->col_ptr, lp->row_idx, A, y, v);
        for (int j = 0; j < n; ++j) {
            const double x_old = x[j];
            const double grad = v[j] + wc[j];
            const double tg = tau * grad;
            double x_new = x_old - tg;
            if (x_new < wcl[j]) x_new = wcl[j];
            if (x_new > wcu[j]) x_new = wcu[j];
            if (isnan(x_new) || isinf(x_new)) has_nan = 1;
            x[j] = x_new;
            const double twice = 2.0 * x_new;
            xb[j] = twice - x_old;
        }
        if (has_nan) { status = ST_NUMERICAL_ERROR; break; }
        if (opts->stop_at_iter > 0 && iter == opts->stop_at_iter) break;

        /* pdhg.cpp:249-365 */
        if (iter % check_freq == 0 || iter == max_iter) {
            const double elapsed = now_sec() - start_time - res->monitor_time;
            if (elapsed > opts->time_limit) { status = ST_TIME_LIMIT; break; }
            or_csc_multiply(m, n, lp->col_ptr, lp->row_idx, A, x, Ax);
            double max_prim = 0.0;
            for (int i = 0; i < m; ++i) {
                const double axi = Ax[i], li = wrl[i], ui = wru[i];
                if (axi < li) max_prim = SMAX(max_prim, li - axi);
                if (axi > ui) max_prim = SMAX(max_prim, axi - ui);
            }
            for (int j = 0; j < n; ++j) {
                const double xj = x[j], lj = wcl[j], uj = wcu[j];
                if (xj < lj) max_prim = SMAX(max_prim, lj - xj);
                if (xj > uj) max_prim = SMAX(max_prim, xj - uj);
            }
            double max_dual = 0.0, dual_obj = 0.0;
            for (int i = 0; i < m; ++i) {
                const double yi = y[i], li = wrl[i], ui = wru[i];
                if (yi > 1e-12 && ui < 1e20) { const double p = yi * ui; dual_obj = dual_obj - p; }
                else if (yi < -1e-12 && li > -1e20) { const double p = yi * li; dual_obj = dual_obj - p; }
            }
            for (int j = 0; j < n; ++j) {
                const double dj = wc[j] + v[j];
                const double xj = x[j], lj = wcl[j], uj = wcu[j];
                if (xj > lj + 1e-5 && dj > 1e-6) max_dual = SMAX(max_dual, dj);
                if (xj < uj - 1e-5 && dj < -1e-6) max_dual = SMAX(max_dual, -dj);
                if (lj <= -1e20 && uj >= 1e20) max_dual = SMAX(max_dual, fabs(dj));
                if (dj > 1e-12 && lj > -1e20) { const double p = dj * lj; dual_obj = dual_obj + p; }
                else if (dj < -1e-12 && uj < 1e20) { const double p = dj * uj; dual_obj = dual_obj + p; }
            }
            double primal_obj = 0.0;
            for (int j = 0; j < n; ++j) { const double p = wc[j] * x[j]; primal_obj = primal_obj + p; }
            const double gap = fabs(primal_obj - dual_obj) / (1.0 + fabs(primal_obj) + fabs(dual_obj));
            res->primal_residual = max_prim;
            res->dual_residual = max_dual;
            res->duality_gap = gap;

            if (mx) {
                const double t0 = now_sec();
                for (int j = 0; j < n; ++j) mx[j] = scaled ? x[j] * cs[j] : x[j];
                for (int i = 0; i < m; ++i) my[i] = -(scaled ? y[i] * rs[i] : y[i]);
                double k[8];
                or_rel_kkt(lp, mx, my, k);
                res->kkt_last = k[0];
                const double t1 = now_sec();
                if (k[0] <= opts->kkt_target && res->kkt_first_iter < 0) {
                    res->kkt_first_iter = iter;
                    res->kkt_first_time = t0 - iter_start - res->monitor_time;
                }
                res->monitor_time += t1 - t0;
            }

            const double total_infeas = max_prim + max_dual;
            if (total_infeas < best_infeas) {
                best_infeas = total_infeas;
                memcpy(best_x, x, sizeof(double) * (size_t)n);
                memcpy(best_y, y, sizeof(double) * (size_t)m);
                stagnation = 0;
            } else {
                stagnation += check_freq;
            }
            if (max_prim <= opts->primal_tolerance && max_dual <= opts->dual_tolerance &&
                gap <= opts->gap_tolerance) {
                converged = 1;
                memcpy(best_x, x, sizeof(double) * (size_t)n);
                memcpy(best_y, y, sizeof(double) * (size_t)m);
                status = ST_OPTIMAL;
                break;
            }
            if (max_prim > 10.0 * max_dual && max_dual > 0.0) {
                sigma = SMIN(sigma * 1.3, 1e6);
                tau = (safety * safety) / (spectral_norm * spectral_norm * sigma);
            } else if (max_dual > 10.0 * max_prim && max_prim > 0.0) {
                tau = SMIN(tau * 1.3, 1e6);
                sigma = (safety * safety) / (spectral_norm * spectral_norm * tau);
            }
            if (stagnation >= opts->restart_stagnation_limit) {
                memcpy(xb, x, sizeof(double) * (size_t)n);
                res->restart_count++;
                stagnation = 0;
            }
        }
    }
    const double iter_end = now_sec();
    res->iteration_time = iter_end - iter_start - res->monitor_time;
    res->total_time = iter_end - start_time - res->monitor_time;
    res->iterations = iter;
    res->converged = converged;
    res->final_tau = tau;
    res->final_sigma = sigma;

    if (traj_x) memcpy(traj_x, x, sizeof(double) * (size_t)n);
    if (traj_xbar) memcpy(traj_xbar, xb, sizeof(double) * (size_t)n);
    if (traj_y) memcpy(traj_y, y, sizeof(double) * (size_t)m);

    /* pdhg.cpp:373-378 */
    if (!converged && status == ST_NOT_SOLVED && iter >= max_iter)
        status = (best_infeas <= 1e-4) ? ST_FEASIBLE : ST_ITERATION_LIMIT;
    res->status = status;

    /* pdhg.cpp:380-418 */
    for (int j = 0; j < n; ++j) x_out[j] = scaled ? best_x[j] * cs[j] : best_x[j];
    double* yraw = dalloc((size_t)m);
    for (int i = 0; i < m; ++i) yraw[i] = scaled ? best_y[i] * rs[i] : best_y[i];
    double* rowv = dalloc((size_t)m);
    or_csc_multiply(m, n, lp->col_ptr, lp->row_idx, lp->vals, x_out, rowv);
    if (row_value_out) memcpy(row_value_out, rowv, sizeof(double) * (size_t)m);
    double obj = lp->offset;
    for (int j = 0; j < n; ++j) { const double p = lp->c[j] * x_out[j]; obj = obj + p; }
    res->objective = obj;
    for (int i = 0; i < m; ++i) row_dual_out[i] = is_max ? yraw[i] : -yraw[i];
    if (col_dual_out) {
        double* aty = dalloc((size_t)n);
        or_csc_multiply_transpose(m, n, lp->col_ptr, lp->row_idx, lp->vals, row_dual_out, aty);
        for (int j = 0; j < n; ++j) col_dual_out[j] = lp->c[j] - aty[j];
        free(aty);
        /* pdhg.cpp:437 -> model.cpp:203-207,303-307: when the status claims a point,
         * recompute_quality overwrites col_dual with the sense-adjusted reduced costs */
        const int claims = status == ST_OPTIMAL || status == ST_FEASIBLE || status == ST_UNBOUNDED ||
                           status == ST_ITERATION_LIMIT || status == ST_TIME_LIMIT;
        if (claims) {
            double q[5];
            or_recompute_quality(lp, x_out, row_dual_out, col_dual_out, q);
        }
    }
    free(yraw); free(rowv);
    free(A); free(wc); free(wcl); free(wcu); free(wrl); free(wru); free(rs); free(cs);
    free(x); free(xb); free(y); free(w); free(v); free(Ax); free(best_x); free(best_y);
    free(mx); free(my);
    return 0;
}

/* model.cpp:203-374, LP branch (no quadratic term, no integers). */
void or_recompute_quality(const or_lp* lp, const double* x, const double* row_dual,
                          double* col_dual_out, double* q) {
    const int m = lp->m, n = lp->n;
    double maxp = 0.0, maxrel = 0.0, maxd = 0.0, maxc = 0.0;
    for (int j = 0; j < n; ++j) {
        const double xj = x[j], lj = lp->col_lower[j], uj = lp->col_upper[j];
        if (xj < lj) {
            const double viol = lj - xj;
            maxp = SMAX(maxp, viol);
            maxrel = SMAX(maxrel, viol / (1.0 + fabs(lj)));
        }
        if (xj > uj) {
            const double viol = xj - uj;
            maxp = SMAX(maxp, viol);
            maxrel = SMAX(maxrel, viol / (1.0 + fabs(uj)));
        }
    }
    double* ax = dalloc((size_t)m);
    or_csc_multiply(m, n, lp->col_ptr, lp->row_idx, lp->vals, x, ax);
    for (int i = 0; i < m; ++i) {
        const double a = ax[i], li = lp->row_lower[i], ui = lp->row_upper[i];
        if (a < li) {
            const double viol = li - a;
            maxp = SMAX(maxp, viol);
            maxrel = SMAX(maxrel, viol / (1.0 + (li > -1e19 ? fabs(li) : 1.0)));
        }
        if (a > ui) {
            const double viol = a - ui;
            maxp = SMAX(maxp, viol);
            maxrel = SMAX(maxrel, viol / (1.0 + (ui < 1e19 ? fabs(ui) : 1.0)));
        }
    }
    free(ax);
    const int primal_ok = (maxp <= kPrimalFeas || maxrel <= kPrimalFeas);
    double* aty = dalloc((size_t)n);
    or_csc_multiply_transpose(m, n, lp->col_ptr, lp->row_idx, lp->vals, row_dual, aty);
    const double sf = lp->maximize ? -1.0 : 1.0;
    for (int j = 0; j < n; ++j) {
        const double c_eff = sf * lp->c[j];
        const double dj = c_eff - aty[j];
        if (col_dual_out) col_dual_out[j] = lp->maximize ? -dj : dj;
        const double xj = x[j], lj = lp->col_lower[j], uj = lp->col_upper[j];
        if (xj < uj - kPrimalFeas && dj < -kDualFeas) maxd = SMAX(maxd, -dj);
        if (xj > lj + kPrimalFeas && dj > kDualFeas) maxd = SMAX(maxd, dj);
        if (dj > kDualFeas) maxc = SMAX(maxc, dj * fabs(xj - lj));
        else if (dj < -kDualFeas) maxc = SMAX(maxc, (-dj) * fabs(xj - uj));
    }
    free(aty);
    for (int i = 0; i < m; ++i) {
        const double yi = row_dual[i], li = lp->row_lower[i], ui = lp->row_upper[i];
        if (li <= -1e20 && ui < 1e20) { if (yi > kDualFeas) maxd = SMAX(maxd, yi); }
        else if (li > -1e20 && ui >= 1e20) { if (yi < -kDualFeas) maxd = SMAX(maxd, -yi); }
    }
    q[0] = maxp; q[1] = maxd; q[2] = maxc;
    q[3] = primal_ok; q[4] = (maxd <= kDualFeas);
}

/* verifier.cpp:25-330, LP branch. */
void or_verify_solution(const or_lp* lp, const double* x, const double* row_dual,
                        double reported_objective, double tol, double* v) {
    const int m = lp->m, n = lp->n;
    int nviol = 0;
    double maxb = 0.0, maxr = 0.0, maxd = 0.0, maxc = 0.0;
    for (int j = 0; j < n; ++j) if (isnan(x[j]) || isinf(x[j])) nviol++;
    if (isnan(reported_objective) || isinf(reported_objective)) nviol++;
    for (int j = 0; j < n; ++j) {
        const double xj = x[j], lj = lp->col_lower[j], uj = lp->col_upper[j];
        if (xj < lj - tol) { maxb = SMAX(maxb, lj - xj); nviol++; }
        if (xj > uj + tol) { maxb = SMAX(maxb, xj - uj); nviol++; }
    }
    const int bounds_ok = maxb <= tol;
    double* ax = dalloc((size_t)m);
    or_csc_multiply(m, n, lp->col_ptr, lp->row_idx, lp->vals, x, ax);
    for (int i = 0; i < m; ++i) {
        const double a = ax[i], li = lp->row_lower[i], ui = lp->row_upper[i];
        if (a < li - tol) { maxr = SMAX(maxr, li - a); nviol++; }
        if (a > ui + tol) { maxr = SMAX(maxr, a - ui); nviol++; }
    }
    free(ax);
    const double maxp = SMAX(maxb, maxr);
    const int primal_ok = maxp <= tol;
    double obj = lp->offset;
    for (int j = 0; j < n; ++j) { const double p = lp->c[j] * x[j]; obj = obj + p; }
    const double disc = fabs(obj - reported_objective);
    const double denom = SMAX(1.0, fabs(reported_objective));
    const int obj_ok = disc / denom <= tol;
    if (!obj_ok) nviol++;
    int dual_ok = 1;
    if (row_dual) {
        double* aty = dalloc((size_t)n);
        or_csc_multiply_transpose(m, n, lp->col_ptr, lp->row_idx, lp->vals, row_dual, aty);
        const double sf = lp->maximize ? -1.0 : 1.0;
        for (int j = 0; j < n; ++j) {
            const double dj = sf * lp->c[j] - aty[j];
            const double xj = x[j], lj = lp->col_lower[j], uj = lp->col_upper[j];
            if (xj < uj - tol && dj < -tol) { maxd = SMAX(maxd, -dj); nviol++; }
            if (xj > lj + tol && dj > tol) { maxd = SMAX(maxd, dj); nviol++; }
            if (dj > tol) maxc = SMAX(maxc, dj * fabs(xj - lj));
            else if (dj < -tol) maxc = SMAX(maxc, (-dj) * fabs(xj - uj));
        }
        free(aty);
        for (int i = 0; i < m; ++i) {
            const double yi = row_dual[i], li = lp->row_lower[i], ui = lp->row_upper[i];
            if (li <= -1e20 && ui < 1e20) { if (yi > tol) { maxd = SMAX(maxd, yi); nviol++; } }
            else if (li > -1e20 && ui >= 1e20) { if (yi < -tol) { maxd = SMAX(maxd, -yi); nviol++; } }
        }
        dual_ok = (maxd <= tol && maxc <= tol);
    }
    v[0] = primal_ok && bounds_ok && obj_ok && dual_ok && nviol == 0;
    v[1] = maxp; v[2] = maxd; v[3] = maxc; v[4] = disc; v[5] = obj;
    v[6] = primal_ok; v[7] = dual_ok; v[8] = bounds_ok; v[9] = obj_ok;
}
