// LP model, solution and the dispatcher seam (restates include/indus/model.hpp:16-121 of the
// reference). Model: min/max c^T x + offset, row_lower <= A x <= row_upper,
// col_lower <= x <= col_upper, A in CSC; infinite bounds are +-1e20 or +-inf.
#pragma once
#include <string>
#include <vector>

#include "indus/options.hpp"
#include "indus/sparse.hpp"
#include "indus/types.hpp"

namespace indus {

enum class ProblemClass : uint8_t { kLp, kMilp, kQp, kMiqp };

class Model {
public:
    std::string name;
    std::string source_path;
    ObjSense sense = ObjSense::kMinimize;
    double objective_offset = 0.0;
    int num_rows = 0;
    int num_cols = 0;
    la::SparseMatrixCSC A;
    std::vector<double> c;
    std::vector<double> row_lower;
    std::vector<double> row_upper;
    std::vector<double> col_lower;
    std::vector<double> col_upper;
    std::vector<VarType> col_type;
    std::vector<std::string> row_names;
    std::vector<std::string> col_names;
    la::SparseMatrixCSC Q;  // quadratic objectives are outside this library's PDHG LP path

    [[nodiscard]] bool has_quadratic_objective() const noexcept { return Q.nnz() > 0; }
    [[nodiscard]] bool has_integers() const noexcept;
    [[nodiscard]] ProblemClass classify() const noexcept;
    // model.cpp:35-97 (LP part): dimensions, NaN bounds, lower > upper; throws invalid_argument
    void validate() const;
};

struct SolutionQuality {
    double max_primal_violation = 0.0;
    double max_dual_violation = 0.0;
    double max_complementarity_violation = 0.0;
    double max_integrality_violation = 0.0;
    double max_stationarity_residual = 0.0;
    double duality_gap = 0.0;
    double relative_duality_gap = 0.0;
    bool is_primal_feasible = false;
    bool is_dual_feasible = false;
    bool is_integer_feasible = true;
    bool is_stationary = false;
};

class Solution {
public:
    SolveStatus status = SolveStatus::kNotSolved;
    std::string status_message;
    double objective_value = 0.0;
    double best_dual_bound = 0.0;
    double relative_gap = 0.0;
    double absolute_gap = 0.0;
    int64_t open_nodes = 0;
    int64_t pruned_nodes = 0;
    bool has_incumbent = false;
    bool search_completed = false;
    std::string termination_reason;
    std::vector<double> col_value;
    std::vector<double> row_value;
    std::vector<double> row_dual;
    std::vector<double> col_dual;
    std::vector<BasisStatus> col_basis_status;
    std::vector<BasisStatus> row_basis_status;
    std::string certificate_type = "none";
    std::vector<double> certificate_vector;
    std::string algorithm_used = "dual_simplex";
    int64_t iterations = 0;
    int64_t nodes = 0;
    double solve_time_seconds = 0.0;
    int presolve_num_rows = -1;
    int presolve_num_cols = -1;
    int presolve_total_reductions = 0;
    int presolve_doubleton_reductions = 0;
    SolutionQuality quality;

    // model.cpp:203-374 (LP branch): FP64 KKT audit against the original model.
    void recompute_quality(const Model& original_model);
};

// The dispatcher seam (model.cpp:376-667) for the PDHG path: validation, trivial models,
// presolve handoff, PDHG dispatch (GPU for "pdhg_cuda" / "pdhg_hip" / use_gpu, CPU for
// "pdhg" / "pdhg_cpu"), quality recomputation and the status guard. Other algorithms
// (simplex, MILP, QP) are not part of this library and return kUnsupported.
Solution solve(const Model& model, const Options& options = Options());

}  // namespace indus
