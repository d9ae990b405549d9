// CPU PDHG: the reference iteration map on the host (semantics of
// src/solvers/pdhg/pdhg.cpp:11-453 of the reference). Serves solve_pdhg_cpu (and
// solve_pdhg_gpu without a device when "allow_cpu_fallback" is set); the GPU engine's
// VJ_MODE_REFERENCE runs the same map.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <iostream>

#include "indus/pdhg.hpp"
#include "indus/scaling.hpp"
#include "indus/tolerances.hpp"

namespace indus::pdhg {

namespace {
using clk = std::chrono::high_resolution_clock;
double since(clk::time_point t) { return std::chrono::duration<double>(clk::now() - t).count(); }
bool finite(double v) { return !std::isnan(v) && !std::isinf(v); }
}  // namespace

double estimate_spectral_norm(const la::SparseMatrixCSC& A, int max_iter, double tol) {
    const int m = A.m, n = A.n;
    if (m == 0 || n == 0 || A.nnz() == 0) return 1.0;
    std::vector<double> v(n, 1.0 / std::sqrt(static_cast<double>(n))), w(m), u(n);
    double lambda_prev = 0.0, lambda = 1.0;
    for (int it = 0; it < max_iter; ++it) {
        A.multiply(v, w);
        A.multiply_transpose(w, u);
        double sq = 0.0;
        for (double e : u) sq += e * e;
        const double norm = std::sqrt(sq);
        if (norm < 1e-15 || !finite(norm)) return lambda > 1e-6 ? std::sqrt(lambda) : 1e-3;
        lambda = norm;
        if (it > 0 && std::fabs(lambda - lambda_prev) <= tol * lambda) break;
        lambda_prev = lambda;
        const double inv = 1.0 / norm;
        for (int j = 0; j < n; ++j) v[j] = u[j] * inv;
    }
    const double s = std::sqrt(lambda);
    return s > 1e-12 ? s : 1e-6;
}

CpuPdhgSolver::CpuPdhgSolver(const Model& model, const Options& options) : original_model_(model), options_(options) {
    pdhg_opts_.max_iterations = options.iteration_limit;
    pdhg_opts_.time_limit = options.time_limit;
    pdhg_opts_.primal_tolerance = options.get_double("tolerance", 1e-6);
    pdhg_opts_.dual_tolerance = options.get_double("tolerance", 1e-6);
    pdhg_opts_.gap_tolerance = options.get_double("gap_tolerance", 1e-6);
    pdhg_opts_.check_frequency = static_cast<int>(options.get_int("check_frequency", 100));
    pdhg_opts_.enable_scaling = options.enable_scaling;
    pdhg_opts_.verbose = options.get_bool("verbose", false);
}

Solution CpuPdhgSolver::solve() {
    const auto t_start = clk::now();
    diag_ = {};
    const Model& M = original_model_;
    const int m = M.num_rows, n = M.num_cols;
    const bool is_max = M.sense == ObjSense::kMaximize;
    Solution sol;
    sol.algorithm_used = "pdhg_cpu";
    sol.col_value.assign(n, 0.0);
    sol.row_value.assign(m, 0.0);
    sol.row_dual.assign(m, 0.0);
    sol.col_dual.assign(n, 0.0);
    sol.col_basis_status.assign(n, BasisStatus::kUnknown);
    sol.row_basis_status.assign(m, BasisStatus::kUnknown);

    if (n == 0) {  // pdhg.cpp:88-94
        sol.status = SolveStatus::kOptimal;
        sol.objective_value = M.objective_offset;
        sol.solve_time_seconds = since(t_start);
        diag_.converged = true;
        return sol;
    }
    if (m == 0) {  // pdhg.cpp:96-127: each column at its cheapest bound
        double obj = M.objective_offset;
        for (int j = 0; j < n; ++j) {
            const double cj = is_max ? -M.c[j] : M.c[j];
            const double lj = M.col_lower[j], uj = M.col_upper[j];
            double xj;
            if (cj > tol::kDualFeasibility) {
                if (lj <= -1e20) {
                    sol.status = SolveStatus::kUnbounded;
                    sol.status_message = "Unbounded column in unconstrained LP";
                    return sol;
                }
                xj = lj;
            } else if (cj < -tol::kDualFeasibility) {
                if (uj >= 1e20) {
                    sol.status = SolveStatus::kUnbounded;
                    sol.status_message = "Unbounded column in unconstrained LP";
                    return sol;
                }
                xj = uj;
            } else {
                xj = lj > 0.0 ? lj : (uj < 0.0 ? uj : 0.0);
            }
            sol.col_value[j] = xj;
            obj += M.c[j] * xj;
        }
        sol.status = SolveStatus::kOptimal;
        sol.objective_value = obj;
        sol.solve_time_seconds = since(t_start);
        diag_.converged = true;
        return sol;
    }

    // working copies, canonical minimisation
    la::SparseMatrixCSC A = M.A;
    std::vector<double> c = M.c;
    if (is_max)
        for (double& v : c) v = -v;
    std::vector<double> rl = M.row_lower, ru = M.row_upper, cl = M.col_lower, cu = M.col_upper;
    const auto t_setup = clk::now();
    la::ScalingFactors sf;
    if (pdhg_opts_.enable_scaling) sf = la::RuizScaler::compute_and_scale(A, c, rl, ru, cl, cu, 10);
    const double L = estimate_spectral_norm(A, 25, 1e-4);
    diag_.spectral_norm_estimate = L;
    const double safety = std::clamp(pdhg_opts_.step_size_safety, 0.5, 0.99);
    double tau = safety / L, sigma = safety / L;
    diag_.tau = tau;
    diag_.sigma = sigma;

    std::vector<double> x(n), xbar(n), y(m, 0.0), w(m), v(n), ax(m);
    for (int j = 0; j < n; ++j) {
        double val = 0.0;
        if (val < cl[j]) val = cl[j];
        if (val > cu[j]) val = cu[j];
        x[j] = xbar[j] = val;
    }
    std::vector<double> best_x = x, best_y = y;
    double best_infeas = 1e20;
    int stagnation = 0;
    diag_.setup_time_sec = since(t_setup);

    const auto t_iter = clk::now();
    bool converged = false;
    const int64_t max_iter = pdhg_opts_.max_iterations;
    const int freq = std::max(1, pdhg_opts_.check_frequency);
    int64_t iter = 1;
    for (; iter <= max_iter; ++iter) {
        // dual step (pdhg.cpp:197-220): Moreau identity for the row-bound indicator
        A.multiply(xbar, w);
        bool bad = false;
        for (int i = 0; i < m; ++i) {
            const double yh = y[i] + sigma * w[i];
            const double val = yh / sigma;
            const double cl_i = val < rl[i] ? rl[i] : (val > ru[i] ? ru[i] : val);
            y[i] = yh - sigma * cl_i;
            if (!finite(y[i])) bad = true;
        }
        if (bad) {
            sol.status = SolveStatus::kNumericalError;
            sol.status_message = "PDHG dual iterate encountered NaN/Inf";
            break;
        }
        // primal step and extrapolation (pdhg.cpp:222-247)
        A.multiply_transpose(y, v);
        for (int j = 0; j < n; ++j) {
            const double x_old = x[j];
            double xn = x_old - tau * (v[j] + c[j]);
            if (xn < cl[j]) xn = cl[j];
            if (xn > cu[j]) xn = cu[j];
            if (!finite(xn)) bad = true;
            x[j] = xn;
            xbar[j] = 2.0 * xn - x_old;
        }
        if (bad) {
            sol.status = SolveStatus::kNumericalError;
            sol.status_message = "PDHG primal iterate encountered NaN/Inf";
            break;
        }
        if (iter % freq != 0 && iter != max_iter) continue;

        // check (pdhg.cpp:249-365)
        if (since(t_start) > pdhg_opts_.time_limit) {
            sol.status = SolveStatus::kTimeLimit;
            sol.status_message = "PDHG time limit reached";
            break;
        }
        A.multiply(x, ax);
        double prim = 0.0;
        for (int i = 0; i < m; ++i) {
            if (ax[i] < rl[i]) prim = std::max(prim, rl[i] - ax[i]);
            if (ax[i] > ru[i]) prim = std::max(prim, ax[i] - ru[i]);
        }
        for (int j = 0; j < n; ++j) {
            if (x[j] < cl[j]) prim = std::max(prim, cl[j] - x[j]);
            if (x[j] > cu[j]) prim = std::max(prim, x[j] - cu[j]);
        }
        double dual = 0.0, dobj = 0.0;
        for (int i = 0; i < m; ++i) {
            if (y[i] > 1e-12 && ru[i] < 1e20) dobj -= y[i] * ru[i];
            else if (y[i] < -1e-12 && rl[i] > -1e20) dobj -= y[i] * rl[i];
        }
        for (int j = 0; j < n; ++j) {
            const double dj = c[j] + v[j];
            if (x[j] > cl[j] + 1e-5 && dj > 1e-6) dual = std::max(dual, dj);
            if (x[j] < cu[j] - 1e-5 && dj < -1e-6) dual = std::max(dual, -dj);
            if (cl[j] <= -1e20 && cu[j] >= 1e20) dual = std::max(dual, std::fabs(dj));
            if (dj > 1e-12 && cl[j] > -1e20) dobj += dj * cl[j];
            else if (dj < -1e-12 && cu[j] < 1e20) dobj += dj * cu[j];
        }
        double pobj = 0.0;
        for (int j = 0; j < n; ++j) pobj += c[j] * x[j];
        const double gap = std::fabs(pobj - dobj) / (1.0 + std::fabs(pobj) + std::fabs(dobj));
        diag_.primal_residual = prim;
        diag_.dual_residual = dual;
        diag_.duality_gap = gap;
        if (prim + dual < best_infeas) {
            best_infeas = prim + dual;
            best_x = x;
            best_y = y;
            stagnation = 0;
        } else {
            stagnation += freq;
        }
        if (prim <= pdhg_opts_.primal_tolerance && dual <= pdhg_opts_.dual_tolerance && gap <= pdhg_opts_.gap_tolerance) {
            converged = true;
            diag_.converged = true;
            best_x = x;
            best_y = y;
            sol.status = SolveStatus::kOptimal;
            sol.status_message = "PDHG converged to optimal primal and dual tolerance";
            break;
        }
        if (prim > 10.0 * dual && dual > 0.0) {
            sigma = std::min(sigma * 1.3, 1e6);
            tau = (safety * safety) / (L * L * sigma);
        } else if (dual > 10.0 * prim && prim > 0.0) {
            tau = std::min(tau * 1.3, 1e6);
            sigma = (safety * safety) / (L * L * tau);
        }
        if (stagnation >= 200) {  // "restart": reset the extrapolation point
            xbar = x;
            diag_.restart_count++;
            stagnation = 0;
        }
    }
    diag_.iteration_time_sec = since(t_iter);
    diag_.total_time_sec = since(t_start);
    diag_.iterations = iter;
    if (!converged && sol.status == SolveStatus::kNotSolved && iter >= max_iter) {
        sol.status = best_infeas <= 1e-4 ? SolveStatus::kFeasible : SolveStatus::kIterationLimit;
        sol.status_message = "PDHG reached iteration limit (" + std::to_string(max_iter) + ")";
    }

    // finalisation (pdhg.cpp:380-437)
    la::RuizScaler::unscale_primal(sf, best_x);
    la::RuizScaler::unscale_dual(sf, best_y);
    sol.col_value = best_x;
    M.A.multiply(sol.col_value, sol.row_value);
    double obj = M.objective_offset;
    for (int j = 0; j < n; ++j) obj += M.c[j] * sol.col_value[j];
    sol.objective_value = obj;
    sol.best_dual_bound = obj;
    // Row duals of the canonical minimisation for both senses (y > 0 prices a lower row
    // bound), the convention recompute_quality and the verifier check against. The reference
    // writes +y for maximisation (pdhg.cpp:409-411), which its own quality check then reads
    // with the opposite sign; this API keeps the duals consistent instead.
    for (int i = 0; i < m; ++i) sol.row_dual[i] = -best_y[i];
    std::vector<double> aty(n);
    M.A.multiply_transpose(sol.row_dual, aty);
    for (int j = 0; j < n; ++j) {
        sol.col_dual[j] = M.c[j] - aty[j];
        const double xj = sol.col_value[j], lj = M.col_lower[j], uj = M.col_upper[j];
        if (std::fabs(lj - uj) <= tol::kZeroDrop) sol.col_basis_status[j] = BasisStatus::kFixed;
        else if (std::fabs(xj - lj) <= 1e-5) sol.col_basis_status[j] = BasisStatus::kAtLower;
        else if (std::fabs(xj - uj) <= 1e-5) sol.col_basis_status[j] = BasisStatus::kAtUpper;
        else sol.col_basis_status[j] = BasisStatus::kBasic;
    }
    sol.iterations = iter;
    sol.solve_time_seconds = diag_.total_time_sec;
    sol.recompute_quality(M);
    if (pdhg_opts_.verbose)
        std::cout << "[pdhg_cpu] " << to_string(sol.status) << " iterations=" << iter << " obj=" << obj << "\n";
    return sol;
}

namespace {
thread_local PdhgDiagnostics g_last_cpu;
}

const PdhgDiagnostics& get_last_cpu_pdhg_diagnostics() noexcept { return g_last_cpu; }

Solution solve_pdhg_cpu(const Model& model, const Options& options) {
    CpuPdhgSolver solver(model, options);
    Solution sol = solver.solve();
    g_last_cpu = solver.diagnostics();
    return sol;
}

}  // namespace indus::pdhg
