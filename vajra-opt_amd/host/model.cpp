// Model checks, solution quality and the dispatcher seam (semantics of
// src/engine/model.cpp:18-97, 203-374, 376-667 of the reference, LP path).
#include "indus/model.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <stdexcept>

#include "indus/gpu.hpp"
#include "indus/pdhg.hpp"
#include "indus/presolve.hpp"
#include "indus/tolerances.hpp"
#include "indus/verifier.hpp"

namespace indus {

bool Model::has_integers() const noexcept {
    return std::any_of(col_type.begin(), col_type.end(), [](VarType t) { return t == VarType::kInteger; });
}

ProblemClass Model::classify() const noexcept {
    const bool q = has_quadratic_objective(), i = has_integers();
    if (q && i) return ProblemClass::kMiqp;
    if (q) return ProblemClass::kQp;
    if (i) return ProblemClass::kMilp;
    return ProblemClass::kLp;
}

void Model::validate() const {
    if (num_rows < 0 || num_cols < 0) throw std::invalid_argument("Model dimensions cannot be negative");
    const bool dims_ok = num_rows == 0 ? (A.m == 0 && (A.n == 0 || A.n == num_cols))
                                       : (A.m == num_rows && A.n == num_cols);
    if (!dims_ok) throw std::invalid_argument("Constraint matrix A dimensions do not match model num_rows/num_cols");
    if (static_cast<int>(c.size()) != num_cols) throw std::invalid_argument("Objective vector c size does not match num_cols");
    if (static_cast<int>(col_lower.size()) != num_cols || static_cast<int>(col_upper.size()) != num_cols)
        throw std::invalid_argument("Column bounds sizes do not match num_cols");
    if (static_cast<int>(row_lower.size()) != num_rows || static_cast<int>(row_upper.size()) != num_rows)
        throw std::invalid_argument("Row bounds sizes do not match num_rows");
    if (!col_type.empty() && static_cast<int>(col_type.size()) != num_cols)
        throw std::invalid_argument("col_type size does not match num_cols");
    for (int j = 0; j < num_cols; ++j) {
        if (std::isnan(col_lower[j]) || std::isnan(col_upper[j]))
            throw std::invalid_argument("Column " + std::to_string(j) + " contains NaN bound");
        if (col_lower[j] > col_upper[j] + tol::kZeroDrop)
            throw std::invalid_argument("Column " + std::to_string(j) + " lower bound exceeds upper bound");
    }
    for (int i = 0; i < num_rows; ++i) {
        if (std::isnan(row_lower[i]) || std::isnan(row_upper[i]))
            throw std::invalid_argument("Row " + std::to_string(i) + " contains NaN bound");
        if (row_lower[i] > row_upper[i] + tol::kZeroDrop)
            throw std::invalid_argument("Row " + std::to_string(i) + " lower bound exceeds upper bound");
    }
}

void Solution::recompute_quality(const Model& M) {
    quality = {};
    if (!claims_a_point(status)) return;
    const int m = M.num_rows, n = M.num_cols;
    if (static_cast<int>(col_value.size()) != n) return;
    double rel = 0.0;
    auto note = [&](double viol, double denom) {
        quality.max_primal_violation = std::max(quality.max_primal_violation, viol);
        rel = std::max(rel, viol / denom);
    };
    for (int j = 0; j < n; ++j) {
        const double xj = col_value[j], lj = M.col_lower[j], uj = M.col_upper[j];
        if (xj < lj) note(lj - xj, 1.0 + std::fabs(lj));
        if (xj > uj) note(xj - uj, 1.0 + std::fabs(uj));
    }
    M.A.multiply(col_value, row_value);
    for (int i = 0; i < m; ++i) {
        const double ax = row_value[i], li = M.row_lower[i], ui = M.row_upper[i];
        if (ax < li) note(li - ax, 1.0 + (li > -1e19 ? std::fabs(li) : 1.0));
        if (ax > ui) note(ax - ui, 1.0 + (ui < 1e19 ? std::fabs(ui) : 1.0));
    }
    quality.is_primal_feasible =
        quality.max_primal_violation <= tol::kPrimalFeasibility || rel <= tol::kPrimalFeasibility;
    quality.is_integer_feasible = true;
    if (static_cast<int>(row_dual.size()) != m) return;
    std::vector<double> aty(n);
    M.A.multiply_transpose(row_dual, aty);
    const bool is_max = M.sense == ObjSense::kMaximize;
    col_dual.resize(n);
    for (int j = 0; j < n; ++j) {
        const double dj = (is_max ? -M.c[j] : M.c[j]) - aty[j];
        col_dual[j] = is_max ? -dj : dj;
        const double xj = col_value[j], lj = M.col_lower[j], uj = M.col_upper[j];
        if (xj < uj - tol::kPrimalFeasibility && dj < -tol::kDualFeasibility)
            quality.max_dual_violation = std::max(quality.max_dual_violation, -dj);
        if (xj > lj + tol::kPrimalFeasibility && dj > tol::kDualFeasibility)
            quality.max_dual_violation = std::max(quality.max_dual_violation, dj);
        if (dj > tol::kDualFeasibility)
            quality.max_complementarity_violation = std::max(quality.max_complementarity_violation, dj * std::fabs(xj - lj));
        else if (dj < -tol::kDualFeasibility)
            quality.max_complementarity_violation =
                std::max(quality.max_complementarity_violation, -dj * std::fabs(xj - uj));
    }
    for (int i = 0; i < m; ++i) {
        const double yi = row_dual[i], li = M.row_lower[i], ui = M.row_upper[i];
        if (li <= -1e20 && ui < 1e20 && yi > tol::kDualFeasibility)
            quality.max_dual_violation = std::max(quality.max_dual_violation, yi);
        else if (li > -1e20 && ui >= 1e20 && yi < -tol::kDualFeasibility)
            quality.max_dual_violation = std::max(quality.max_dual_violation, -yi);
    }
    quality.is_dual_feasible = quality.max_dual_violation <= tol::kDualFeasibility;
}

namespace {

using clk = std::chrono::high_resolution_clock;
double since(clk::time_point t) { return std::chrono::duration<double>(clk::now() - t).count(); }

Solution solve_trivial_zero_cols(const Model& model) {  // model.cpp:411-427
    Solution sol;
    sol.objective_value = model.objective_offset;
    for (int i = 0; i < model.num_rows; ++i)
        if (0.0 < model.row_lower[i] - tol::kPrimalFeasibility || 0.0 > model.row_upper[i] + tol::kPrimalFeasibility) {
            sol.status = SolveStatus::kInfeasible;
            sol.status_message = "Empty model constraint violated at row " + std::to_string(i);
            return sol;
        }
    sol.status = SolveStatus::kOptimal;
    sol.status_message = "Model with zero variables solved trivially";
    return sol;
}

Solution solve_trivial_zero_rows(const Model& model) {  // model.cpp:429-469
    Solution sol;
    sol.col_value.assign(model.num_cols, 0.0);
    sol.col_dual.assign(model.num_cols, 0.0);
    double obj = model.objective_offset;
    for (int j = 0; j < model.num_cols; ++j) {
        const double raw = model.c[j];
        const double cj = model.sense == ObjSense::kMaximize ? -raw : raw;
        const double lj = model.col_lower[j], uj = model.col_upper[j];
        double xj;
        if (cj > tol::kDualFeasibility) {
            if (lj <= -1e20) {
                sol.status = SolveStatus::kUnbounded;
                sol.status_message = "Unbounded variable " + std::to_string(j);
                return sol;
            }
            xj = lj;
        } else if (cj < -tol::kDualFeasibility) {
            if (uj >= 1e20) {
                sol.status = SolveStatus::kUnbounded;
                sol.status_message = "Unbounded variable " + std::to_string(j);
                return sol;
            }
            xj = uj;
        } else {
            xj = std::clamp(0.0, lj, uj);
        }
        sol.col_value[j] = xj;
        sol.col_dual[j] = cj;
        obj += raw * xj;
    }
    sol.status = SolveStatus::kOptimal;
    sol.status_message = "Unconstrained model solved trivially";
    sol.objective_value = obj;
    return sol;
}

bool wants_gpu(const Options& o) {
    return o.algorithm == "pdhg_cuda" || o.algorithm == "pdhg_hip" || o.use_gpu;
}
bool wants_cpu_pdhg(const Options& o) { return o.algorithm == "pdhg" || o.algorithm == "pdhg_cpu"; }

}  // namespace

Solution solve(const Model& model, const Options& options) {
    const auto t0 = clk::now();
    try {
        model.validate();
    } catch (const std::exception& e) {
        Solution sol;
        sol.status = SolveStatus::kModelError;
        sol.status_message = std::string("Model validation error: ") + e.what();
        sol.solve_time_seconds = since(t0);
        return sol;
    }
    if (model.classify() != ProblemClass::kLp || (!wants_gpu(options) && !wants_cpu_pdhg(options))) {
        Solution sol;
        sol.status = SolveStatus::kUnsupported;
        sol.status_message = "this library provides the PDHG LP path only (algorithm pdhg / pdhg_cpu / pdhg_cuda / pdhg_hip)";
        sol.solve_time_seconds = since(t0);
        return sol;
    }
    if (options.time_limit <= 0.0) {
        Solution sol;
        sol.status = SolveStatus::kTimeLimit;
        sol.status_message = "Time limit exceeded before solve commenced";
        return sol;
    }
    if (model.num_cols == 0) {
        Solution sol = solve_trivial_zero_cols(model);
        sol.solve_time_seconds = since(t0);
        return sol;
    }
    if (model.num_rows == 0) {
        Solution sol = solve_trivial_zero_rows(model);
        sol.solve_time_seconds = since(t0);
        return sol;
    }

    // presolve -> solve reduced -> postsolve (model.cpp:471-521)
    if (options.enable_presolve) {
        auto pre = presolve::PresolveEngine::apply(model);
        if (pre.is_infeasible || pre.is_unbounded) {
            Solution sol;
            sol.status = pre.is_infeasible ? SolveStatus::kInfeasible : SolveStatus::kUnbounded;
            sol.status_message = std::string(pre.is_infeasible ? "Presolve detected infeasibility: "
                                                               : "Presolve detected unboundedness: ") +
                                 pre.status_message;
            sol.solve_time_seconds = since(t0);
            return sol;
        }
        if (pre.total_reductions() > 0) {
            Options inner = options;
            inner.enable_presolve = false;
            Solution red;
            if (pre.presolved_model.num_cols == 0 || pre.presolved_model.num_rows == 0) {
                // presolve left nothing to solve: the reduced answer is its offset alone
                red.status = SolveStatus::kOptimal;
                red.status_message = "Solved to optimality by presolve reductions";
                red.objective_value = pre.presolved_model.objective_offset;
            } else {
                red = solve(pre.presolved_model, inner);
            }
            Solution full = presolve::PresolveEngine::postsolve(red, pre.stack, model);
            full.solve_time_seconds = since(t0);
            full.presolve_num_rows = pre.presolved_model.num_rows;
            full.presolve_num_cols = pre.presolved_model.num_cols;
            full.presolve_total_reductions = pre.total_reductions();
            full.presolve_doubleton_reductions = pre.num_doubleton_rows;
            if (full.status == SolveStatus::kOptimal) {
                auto v = verifier::verify_solution(model, full, 1e-4);
                if (!v.passed) {
                    full.status = SolveStatus::kNumericalError;
                    full.status_message = "Postsolve verification failed: " + v.summary;
                }
            }
            return full;
        }
    }

    Solution sol = wants_gpu(options) ? gpu::solve_pdhg_gpu(model, options) : pdhg::solve_pdhg_cpu(model, options);
    sol.solve_time_seconds = since(t0);
    sol.presolve_num_rows = model.num_rows;
    sol.presolve_num_cols = model.num_cols;
    sol.recompute_quality(model);

    // status guard (model.cpp:626-664)
    auto bad = [](double v) { return std::isnan(v) || std::isinf(v); };
    if (bad(sol.objective_value)) {
        sol.status = SolveStatus::kNumericalError;
        sol.status_message = "Status downgraded to NUMERICAL_ERROR: objective is NaN or Inf";
    }
    if (std::any_of(sol.col_value.begin(), sol.col_value.end(), bad)) {
        sol.status = SolveStatus::kNumericalError;
        sol.status_message = "Status downgraded to NUMERICAL_ERROR: primal variable is NaN or Inf";
    }
    if (std::any_of(sol.row_dual.begin(), sol.row_dual.end(), bad)) {
        sol.status = SolveStatus::kNumericalError;
        sol.status_message = "Status downgraded to NUMERICAL_ERROR: dual multiplier is NaN or Inf";
    }
    if (sol.status == SolveStatus::kOptimal || sol.status == SolveStatus::kFeasible) {
        if (!sol.quality.is_primal_feasible) {
            sol.status = SolveStatus::kNumericalError;
            sol.status_message = "Status downgraded to NUMERICAL_ERROR: primal feasibility violation (" +
                                 std::to_string(sol.quality.max_primal_violation) + ") exceeds tolerance";
        } else if (sol.status == SolveStatus::kOptimal && !sol.quality.is_dual_feasible) {
            sol.status = SolveStatus::kFeasible;
            sol.status_message = "Status downgraded to FEASIBLE: dual feasibility / complementarity violation (" +
                                 std::to_string(sol.quality.max_dual_violation) + ") exceeds tolerance";
        }
    }
    return sol;
}

}  // namespace indus
