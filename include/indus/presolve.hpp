// Presolve -> PDHG -> postsolve handoff (the contract of include/indus/presolve.hpp:11-108
// of the reference: reversible reductions pushed on a LIFO stack with index maps between the
// reduced and the original model, and a postsolve that rebuilds primal and dual values).
//
// All eight reductions of the reference are applied, in its order within a pass (empty rows,
// empty columns, fixed columns, singleton rows, redundant and forcing rows, free column
// singletons, doubleton equality rows), to a fixed point or max_passes passes, with its tests
// and tolerances; the reduced model, the stack and the postsolved solution are those of the
// reference (src/presolve/presolve.cpp:10-976). Kept reference behaviours: an empty column's
// bound is chosen by its raw cost, also for maximisation; a row touched by a doubleton
// substitution is never a doubleton candidate again in the same call.
//
// Dual convention: duals are those of the canonical minimisation (c' = sense * c), with
// y_i > 0 pricing the row's lower bound and d = c' - A^T y.
#pragma once
#include <string>
#include <utility>
#include <vector>

#include "indus/model.hpp"

namespace indus::presolve {

enum class ReductionType : uint8_t {
    kEmptyRow,
    kEmptyCol,
    kFixedCol,
    kSingletonRow,
    kForcingRow,
    kRedundantRow,
    kFreeColSingleton,
    kDoubletonRow
};

struct ReductionRecord {
    ReductionType type = ReductionType::kEmptyRow;
    int row_idx = -1;
    int col_idx = -1;
    int other_col_idx = -1;
    std::string name;
    double coeff = 0.0;
    double other_coeff = 0.0;
    double old_lower = 0.0;   // column bounds before a singleton row tightened them
    double old_upper = 0.0;
    double row_lower = 0.0;
    double row_upper = 0.0;
    double fixed_value = 0.0;
    double obj_cost = 0.0;
    std::vector<std::pair<int, double>> entries;
};

class PresolveStack {
public:
    void push(ReductionRecord rec) { records_.push_back(std::move(rec)); }
    [[nodiscard]] size_t size() const noexcept { return records_.size(); }
    [[nodiscard]] bool empty() const noexcept { return records_.empty(); }
    [[nodiscard]] const std::vector<ReductionRecord>& records() const noexcept { return records_; }
    std::vector<ReductionRecord>& records() noexcept { return records_; }

    std::vector<int> reduced_to_orig_row;
    std::vector<int> reduced_to_orig_col;
    std::vector<int> orig_to_reduced_row;
    std::vector<int> orig_to_reduced_col;

private:
    std::vector<ReductionRecord> records_;
};

struct PresolveResult {
    Model presolved_model;
    PresolveStack stack;
    bool is_infeasible = false;
    bool is_unbounded = false;
    std::string status_message;
    int num_empty_rows = 0;
    int num_empty_cols = 0;
    int num_fixed_cols = 0;
    int num_singleton_rows = 0;
    int num_forcing_rows = 0;
    int num_redundant_rows = 0;
    int num_free_col_singletons = 0;
    int num_doubleton_rows = 0;
    [[nodiscard]] int total_reductions() const noexcept {
        return num_empty_rows + num_empty_cols + num_fixed_cols + num_singleton_rows + num_forcing_rows +
               num_redundant_rows + num_free_col_singletons + num_doubleton_rows;
    }
};

class PresolveEngine {
public:
    static PresolveResult apply(const Model& original_model, int max_passes = 10);
    static Solution postsolve(const Solution& reduced_solution, const PresolveStack& stack,
                              const Model& original_model);
    // col_dual = reduced costs of the row duals on the original model (tol is unused)
    static void refine_dual_solution(const Model& original_model, Solution& solution, double tol = 1e-6);
};

}  // namespace indus::presolve
