// TEST INFRASTRUCTURE -- a canonical text dump of a presolve result and of the postsolve of a
// fixed synthetic reduced solution. Written against the field names the reference's
// indus::presolve / indus::Model / indus::Solution types share with this repository's, it is
// compiled twice: into the oracle's shim (reference types, oracle/ref_shim.cpp) and into the
// C++ test driver (this repository's types, tests/cpp/test_indus_api.cpp). Equal dumps mean the
// same reductions, in the same order, with bit-identical numbers (doubles printed with %a).
#pragma once
#include <cstdio>
#include <string>
#include <vector>

namespace presolve_dump {

inline void num(FILE* f, const char* k, double v) { std::fprintf(f, "%s %a\n", k, v); }

template <class Vec>
inline void vec(FILE* f, const char* k, const Vec& v) {
    std::fprintf(f, "%s %zu", k, v.size());
    for (const auto& x : v) std::fprintf(f, " %a", static_cast<double>(x));
    std::fprintf(f, "\n");
}

template <class Vec>
inline void ivec(FILE* f, const char* k, const Vec& v) {
    std::fprintf(f, "%s %zu", k, v.size());
    for (const auto& x : v) std::fprintf(f, " %d", static_cast<int>(x));
    std::fprintf(f, "\n");
}

// The presolve result: status, counts, reduced model, index maps, reduction records.
template <class Result>
void dump_result(FILE* f, const Result& r) {
    std::fprintf(f, "infeasible %d unbounded %d\nmessage %s\n", r.is_infeasible ? 1 : 0, r.is_unbounded ? 1 : 0,
                 r.status_message.c_str());
    std::fprintf(f, "counts %d %d %d %d %d %d %d %d total %d\n", r.num_empty_rows, r.num_empty_cols, r.num_fixed_cols,
                 r.num_singleton_rows, r.num_forcing_rows, r.num_redundant_rows, r.num_free_col_singletons,
                 r.num_doubleton_rows, r.total_reductions());
    if (r.is_infeasible || r.is_unbounded) return;
    const auto& P = r.presolved_model;
    std::fprintf(f, "reduced %d %d nnz %zu name %s max %d\n", P.num_rows, P.num_cols, P.A.row_idx.size(),
                 P.name.c_str(), static_cast<int>(P.sense));
    num(f, "offset", P.objective_offset);
    vec(f, "c", P.c);
    vec(f, "col_lower", P.col_lower);
    vec(f, "col_upper", P.col_upper);
    vec(f, "row_lower", P.row_lower);
    vec(f, "row_upper", P.row_upper);
    ivec(f, "col_ptr", P.A.col_ptr);
    ivec(f, "row_idx", P.A.row_idx);
    vec(f, "values", P.A.values);
    for (const auto& s : P.col_names) std::fprintf(f, "col_name %s\n", s.c_str());
    for (const auto& s : P.row_names) std::fprintf(f, "row_name %s\n", s.c_str());
    const auto& S = r.stack;
    ivec(f, "red_to_orig_row", S.reduced_to_orig_row);
    ivec(f, "red_to_orig_col", S.reduced_to_orig_col);
    ivec(f, "orig_to_red_row", S.orig_to_reduced_row);
    ivec(f, "orig_to_red_col", S.orig_to_reduced_col);
    for (const auto& q : S.records()) {
        std::fprintf(f, "record %d row %d col %d other %d name %s\n", static_cast<int>(q.type), q.row_idx, q.col_idx,
                     q.other_col_idx, q.name.c_str());
        std::fprintf(f, "  %a %a %a %a %a %a %a %a\n", q.coeff, q.other_coeff, q.old_lower, q.old_upper, q.row_lower,
                     q.row_upper, q.fixed_value, q.obj_cost);
        std::fprintf(f, "  entries %zu", q.entries.size());
        for (const auto& e : q.entries) std::fprintf(f, " %d:%a", e.first, e.second);
        std::fprintf(f, "\n");
    }
}

// The synthetic reduced solution both sides postsolve: a deterministic point inside the
// reduced column bounds and deterministic row duals of both signs.
template <class Solution, class Model, class Status>
Solution synthetic_reduced(const Model& P, Status optimal) {
    Solution s;
    s.status = optimal;
    s.status_message = "synthetic";
    s.iterations = 7;
    s.col_value.resize(P.num_cols);
    s.col_dual.assign(P.num_cols, 0.0);
    s.row_dual.resize(P.num_rows);
    s.row_value.assign(P.num_rows, 0.0);
    for (int j = 0; j < P.num_cols; ++j) {
        double v = static_cast<double>((j * 7919) % 1000) / 250.0 - 1.0;
        if (v < P.col_lower[j]) v = P.col_lower[j];
        if (v > P.col_upper[j]) v = P.col_upper[j];
        // every third column sits on a finite bound, so the bound tests of postsolve are used
        if (j % 3 == 0 && P.col_lower[j] > -1e19) v = P.col_lower[j];
        if (j % 3 == 1 && P.col_upper[j] < 1e19) v = P.col_upper[j];
        s.col_value[j] = v;
    }
    for (int i = 0; i < P.num_rows; ++i) s.row_dual[i] = static_cast<double>((i * 104729) % 2001 - 1000) / 700.0;
    return s;
}

template <class Solution>
void dump_solution(FILE* f, const Solution& s) {
    std::fprintf(f, "status %d message %s iterations %lld algorithm %s\n", static_cast<int>(s.status),
                 s.status_message.c_str(), static_cast<long long>(s.iterations), s.algorithm_used.c_str());
    num(f, "objective", s.objective_value);
    num(f, "best_dual_bound", s.best_dual_bound);
    vec(f, "col_value", s.col_value);
    vec(f, "col_dual", s.col_dual);
    vec(f, "row_value", s.row_value);
    vec(f, "row_dual", s.row_dual);
    ivec(f, "col_basis", s.col_basis_status);
    ivec(f, "row_basis", s.row_basis_status);
    const auto& q = s.quality;
    std::fprintf(f, "quality %a %a %a %d %d\n", q.max_primal_violation, q.max_dual_violation,
                 q.max_complementarity_violation, q.is_primal_feasible ? 1 : 0, q.is_dual_feasible ? 1 : 0);
}

}  // namespace presolve_dump
