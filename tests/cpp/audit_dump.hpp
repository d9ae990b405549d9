// Test infrastructure: the host KKT audits of a caller-given point, written as text.
// Compiled against this build's headers (tests/cpp/test_indus_api.cpp, mode audit_dump) and
// against the reference's (oracle/ref_shim.cpp, ref_audit_dump), so the dumps of
// Solution::recompute_quality and verifier::verify_solution can be compared text for text.
//
// Point file (text): "status <int>", "objective <hex double>", "x <hex doubles>",
// "row_dual <hex doubles>" (an empty row_dual line: no duals given).
// Dump: verify_solution on the point as given; recompute_quality (every SolutionQuality
// field, row_value, col_dual); verify_solution again on the audited solution.
#pragma once
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

namespace audit_dump {

struct Point {
    int status = 1;
    double objective = 0.0;
    std::vector<double> x, row_dual;
};

inline bool read_point(const char* path, Point& p) {
    std::ifstream f(path);
    if (!f) return false;
    std::string line;
    while (std::getline(f, line)) {
        std::istringstream in(line);
        std::string key, tok;
        in >> key;
        if (key == "status") {
            in >> p.status;
        } else if (key == "objective") {
            in >> tok;
            p.objective = std::strtod(tok.c_str(), nullptr);
        } else if (key == "x" || key == "row_dual") {
            std::vector<double>& v = key == "x" ? p.x : p.row_dual;
            while (in >> tok) v.push_back(std::strtod(tok.c_str(), nullptr));
        }
    }
    return true;
}

inline void vec(FILE* f, const char* tag, const std::vector<double>& v) {
    std::fprintf(f, "%s %zu", tag, v.size());
    for (double a : v) std::fprintf(f, " %a", a);
    std::fprintf(f, "\n");
}

template <class R>
void verification(FILE* f, const char* tag, const R& r) {
    std::fprintf(f, "%s flags %d %d %d %d %d %d %d %d %d %d\n", tag, r.passed ? 1 : 0, r.primal_feasible ? 1 : 0,
                 r.dual_feasible ? 1 : 0, r.bounds_feasible ? 1 : 0, r.objective_matches ? 1 : 0, r.is_milp ? 1 : 0,
                 r.integer_feasible ? 1 : 0, r.optimality_proven ? 1 : 0, r.is_qp ? 1 : 0, r.is_convex ? 1 : 0);
    std::fprintf(f, "%s values %a %a %a %a %a %a %a %a %a %a %a\n", tag, r.max_integrality_violation, r.mip_gap,
                 r.max_stationarity_residual, r.max_primal_violation, r.max_bound_violation, r.max_row_violation,
                 r.max_dual_violation, r.max_complementarity_violation, r.objective_discrepancy,
                 r.reported_objective, r.recomputed_objective);
    std::fprintf(f, "%s convexity [%s] status [%s] summary [%s]\n", tag, r.convexity_status.c_str(),
                 r.reported_status.c_str(), r.summary.c_str());
    for (const auto& v : r.violations) std::fprintf(f, "%s violation [%s]\n", tag, v.c_str());
}

// Solution: the Solution class of either side; verify(model, solution, tol) its verifier.
template <class Solution, class Model, class Status, class Verify>
void run(FILE* f, const Model& model, const Point& p, double tol, Verify verify) {
    Solution s;
    s.status = static_cast<Status>(p.status);
    s.objective_value = p.objective;
    s.col_value = p.x;
    s.row_dual = p.row_dual;
    verification(f, "given", verify(model, s, tol));
    s.recompute_quality(model);
    const auto& q = s.quality;
    std::fprintf(f, "quality %a %a %a %a %a %a %a %d %d %d %d\n", q.max_primal_violation, q.max_dual_violation,
                 q.max_complementarity_violation, q.max_integrality_violation, q.max_stationarity_residual,
                 q.duality_gap, q.relative_duality_gap, q.is_primal_feasible ? 1 : 0, q.is_dual_feasible ? 1 : 0,
                 q.is_integer_feasible ? 1 : 0, q.is_stationary ? 1 : 0);
    vec(f, "row_value", s.row_value);
    vec(f, "col_dual", s.col_dual);
    verification(f, "audited", verify(model, s, tol));
}

// The Solution both sides' writers are given for a point: what a device PDHG solve reports
// (status, objective and dual bound, x, row duals, algorithm, iterations) plus the row
// activities, reduced costs and quality that recompute_quality fills in.
template <class Solution, class Model, class Status>
Solution solution_of(const Model& model, const Point& p) {
    Solution s;
    s.status = static_cast<Status>(p.status);
    s.objective_value = p.objective;
    s.best_dual_bound = p.objective;
    s.col_value = p.x;
    s.row_dual = p.row_dual;
    s.algorithm_used = "pdhg_cuda";
    s.iterations = 4096;
    s.status_message = "converged";
    s.recompute_quality(model);
    return s;
}

}  // namespace audit_dump
