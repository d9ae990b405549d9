// In-memory KKT verifier (restates include/indus/verifier.hpp:9-49 and the LP branch of
// src/engine/verifier.cpp:25-330 of the reference). File-based verification (.sol parsing)
// is outside this library's PDHG path.
#pragma once
#include <string>
#include <vector>

#include "indus/model.hpp"

namespace indus::verifier {

struct VerificationResult {
    bool passed = false;
    bool primal_feasible = false;
    bool dual_feasible = false;
    bool bounds_feasible = false;
    bool objective_matches = false;
    bool is_milp = false;
    bool integer_feasible = true;
    double max_integrality_violation = 0.0;
    bool optimality_proven = false;
    double mip_gap = 0.0;
    bool is_qp = false;
    bool is_convex = false;
    std::string convexity_status;
    double max_stationarity_residual = 0.0;
    double max_primal_violation = 0.0;
    double max_bound_violation = 0.0;
    double max_row_violation = 0.0;
    double max_dual_violation = 0.0;
    double max_complementarity_violation = 0.0;
    double objective_discrepancy = 0.0;
    double reported_objective = 0.0;
    double recomputed_objective = 0.0;
    std::string reported_status;
    std::vector<std::string> violations;
    std::string summary;
};

VerificationResult verify_solution(const Model& model, const Solution& solution, double tol = 1e-6);

// Build-defined relative KKT error of a solution (FP64, original data), the stopping measure
// of the restarted engine: max of |r_p|/(1+|b|), |r_d|/(1+|c|), |p-d|/(1+|p|+|d|).
struct RelKkt {
    double rel_kkt = 0.0, rel_primal = 0.0, rel_dual = 0.0, rel_gap = 0.0;
    double primal_objective = 0.0, dual_objective = 0.0;
};
RelKkt relative_kkt(const Model& model, const Solution& solution);

}  // namespace indus::verifier
