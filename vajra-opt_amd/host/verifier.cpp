// In-memory verifier (semantics of the LP branch of src/engine/verifier.cpp:25-330 of the
// reference) and the build-defined relative KKT error of a solution.
#include "indus/verifier.hpp"

#include <algorithm>
#include <cmath>
#include <sstream>

namespace indus::verifier {

namespace {
std::string name_of(const std::vector<std::string>& names, int k, char prefix) {
    if (static_cast<size_t>(k) < names.size() && !names[k].empty()) return names[k];
    return std::string(1, prefix) + std::to_string(k);
}
// Values in violation messages: fixed notation, six decimals (the report format the
// reference's verifier writes).
std::string num(double v) { return std::to_string(v); }
}  // namespace

VerificationResult verify_solution(const Model& M, const Solution& sol, double tol) {
    VerificationResult res;
    res.reported_status = sol.status == SolveStatus::kOptimal      ? "OPTIMAL"
                          : sol.status == SolveStatus::kFeasible   ? "FEASIBLE"
                          : sol.status == SolveStatus::kInfeasible ? "INFEASIBLE"
                          : sol.status == SolveStatus::kUnbounded  ? "UNBOUNDED"
                                                                   : "OTHER";
    res.reported_objective = sol.objective_value;
    const int m = M.num_rows, n = M.num_cols;
    if (static_cast<int>(sol.col_value.size()) != n) {
        res.violations.push_back("Primal solution vector size (" + std::to_string(sol.col_value.size()) +
                                 ") does not match model num_cols (" + std::to_string(n) + ")");
        res.summary = "FAIL: Dimension mismatch";
        return res;
    }
    for (int j = 0; j < n; ++j)
        if (std::isnan(sol.col_value[j]) || std::isinf(sol.col_value[j]))
            res.violations.push_back("Variable " + name_of(M.col_names, j, 'c') + " is NaN or Inf");
    if (std::isnan(sol.objective_value) || std::isinf(sol.objective_value))
        res.violations.push_back("Reported objective is NaN or Inf");
    for (int j = 0; j < n; ++j) {
        const double xj = sol.col_value[j], lj = M.col_lower[j], uj = M.col_upper[j];
        if (xj < lj - tol) {
            res.max_bound_violation = std::max(res.max_bound_violation, lj - xj);
            res.violations.push_back("Variable " + name_of(M.col_names, j, 'c') + " below lower bound: " + num(xj) +
                                     " < " + num(lj));
        }
        if (xj > uj + tol) {
            res.max_bound_violation = std::max(res.max_bound_violation, xj - uj);
            res.violations.push_back("Variable " + name_of(M.col_names, j, 'c') + " above upper bound: " + num(xj) +
                                     " > " + num(uj));
        }
    }
    res.bounds_feasible = res.max_bound_violation <= tol;
    std::vector<double> ax(m);
    M.A.multiply(sol.col_value, ax);
    for (int i = 0; i < m; ++i) {
        if (ax[i] < M.row_lower[i] - tol) {
            res.max_row_violation = std::max(res.max_row_violation, M.row_lower[i] - ax[i]);
            res.violations.push_back("Row " + name_of(M.row_names, i, 'r') + " below lower bound: " + num(ax[i]) +
                                     " < " + num(M.row_lower[i]));
        }
        if (ax[i] > M.row_upper[i] + tol) {
            res.max_row_violation = std::max(res.max_row_violation, ax[i] - M.row_upper[i]);
            res.violations.push_back("Row " + name_of(M.row_names, i, 'r') + " above upper bound: " + num(ax[i]) +
                                     " > " + num(M.row_upper[i]));
        }
    }
    res.max_primal_violation = std::max(res.max_bound_violation, res.max_row_violation);
    res.primal_feasible = res.max_primal_violation <= tol;
    double obj = M.objective_offset;
    for (int j = 0; j < n; ++j) obj += M.c[j] * sol.col_value[j];
    res.recomputed_objective = obj;
    res.objective_discrepancy = std::fabs(obj - sol.objective_value);
    res.objective_matches = res.objective_discrepancy / std::max(1.0, std::fabs(sol.objective_value)) <= tol;
    if (!res.objective_matches)
        res.violations.push_back("Objective mismatch: reported " + num(res.reported_objective) + " vs recomputed " +
                                 num(res.recomputed_objective));
    if (!M.has_integers() && static_cast<int>(sol.row_dual.size()) == m) {
        std::vector<double> aty(n);
        M.A.multiply_transpose(sol.row_dual, aty);
        const double sf = M.sense == ObjSense::kMaximize ? -1.0 : 1.0;
        for (int j = 0; j < n; ++j) {
            const double dj = sf * M.c[j] - aty[j];
            const double xj = sol.col_value[j], lj = M.col_lower[j], uj = M.col_upper[j];
            if (xj < uj - tol && dj < -tol) {
                res.max_dual_violation = std::max(res.max_dual_violation, -dj);
                res.violations.push_back("Variable " + name_of(M.col_names, j, 'c') + " can increase but dj < 0: " + num(dj));
            }
            if (xj > lj + tol && dj > tol) {
                res.max_dual_violation = std::max(res.max_dual_violation, dj);
                res.violations.push_back("Variable " + name_of(M.col_names, j, 'c') + " can decrease but dj > 0: " + num(dj));
            }
            if (dj > tol) res.max_complementarity_violation = std::max(res.max_complementarity_violation, dj * std::fabs(xj - lj));
            else if (dj < -tol)
                res.max_complementarity_violation = std::max(res.max_complementarity_violation, -dj * std::fabs(xj - uj));
        }
        for (int i = 0; i < m; ++i) {
            const double yi = sol.row_dual[i], li = M.row_lower[i], ui = M.row_upper[i];
            if (li <= -1e20 && ui < 1e20 && yi > tol) {
                res.max_dual_violation = std::max(res.max_dual_violation, yi);
                res.violations.push_back("Row " + name_of(M.row_names, i, 'r') +
                                         " (<= row) has positive dual multiplier yi: " + num(yi));
            } else if (li > -1e20 && ui >= 1e20 && yi < -tol) {
                res.max_dual_violation = std::max(res.max_dual_violation, -yi);
                res.violations.push_back("Row " + name_of(M.row_names, i, 'r') +
                                         " (>= row) has negative dual multiplier yi: " + num(yi));
            }
        }
        res.dual_feasible = res.max_dual_violation <= tol && res.max_complementarity_violation <= tol;
    } else {
        res.dual_feasible = true;
    }
    res.passed = res.primal_feasible && res.bounds_feasible && res.objective_matches && res.dual_feasible &&
                 res.violations.empty();
    res.optimality_proven = res.passed && sol.status == SolveStatus::kOptimal;
    res.mip_gap = sol.relative_gap;
    std::ostringstream ss;
    ss << (res.passed ? "[VERIFIED OPTIMAL]" : "[VERIFICATION FAILED]") << " Primal viol: " << res.max_primal_violation
       << ", Bound viol: " << res.max_bound_violation << ", Dual viol: " << res.max_dual_violation
       << ", Obj discrepancy: " << res.objective_discrepancy;
    res.summary = ss.str();
    return res;
}

RelKkt relative_kkt(const Model& M, const Solution& sol) {
    RelKkt k;
    const int m = M.num_rows, n = M.num_cols;
    if (static_cast<int>(sol.col_value.size()) != n || static_cast<int>(sol.row_dual.size()) != m) return k;
    const double sense = M.sense == ObjSense::kMaximize ? -1.0 : 1.0;
    std::vector<double> ax(m, 0.0), aty(n, 0.0);
    for (int j = 0; j < n; ++j)
        for (int64_t p = M.A.col_ptr[j]; p < M.A.col_ptr[j + 1]; ++p) {
            ax[M.A.row_idx[p]] += M.A.values[p] * sol.col_value[j];
            aty[j] += M.A.values[p] * sol.row_dual[M.A.row_idx[p]];
        }
    double rp2 = 0, rd2 = 0, b2 = 0, c2 = 0, p = sense * M.objective_offset, d = p;
    for (int i = 0; i < m; ++i) {
        const double l = M.row_lower[i], u = M.row_upper[i], y = sol.row_dual[i];
        const bool lf = l > -1e20, uf = u < 1e20;
        double r = 0;
        if (lf && ax[i] < l) r = ax[i] - l;
        if (uf && ax[i] > u) r = ax[i] - u;
        rp2 += r * r;
        double bm = lf ? std::fabs(l) : 0.0;
        if (uf) bm = std::max(bm, std::fabs(u));
        b2 += bm * bm;
        if (y > 0) { if (lf) d += y * l; else rd2 += y * y; }
        else if (y < 0) { if (uf) d += y * u; else rd2 += y * y; }
    }
    for (int j = 0; j < n; ++j) {
        const double cj = sense * M.c[j], dj = cj - aty[j];
        const double l = M.col_lower[j], u = M.col_upper[j];
        c2 += cj * cj;
        p += cj * sol.col_value[j];
        double lam = 0;
        if (dj > 0 && l > -1e20) lam = dj;
        else if (dj < 0 && u < 1e20) lam = dj;
        rd2 += (dj - lam) * (dj - lam);
        if (lam > 0) d += lam * l;
        else if (lam < 0) d += lam * u;
    }
    k.rel_primal = std::sqrt(rp2) / (1 + std::sqrt(b2));
    k.rel_dual = std::sqrt(rd2) / (1 + std::sqrt(c2));
    k.rel_gap = std::fabs(p - d) / (1 + std::fabs(p) + std::fabs(d));
    k.rel_kkt = std::max({k.rel_primal, k.rel_dual, k.rel_gap});
    k.primal_objective = p;
    k.dual_objective = d;
    return k;
}

}  // namespace indus::verifier
