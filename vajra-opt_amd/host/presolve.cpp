// Presolve and postsolve for the PDHG path (contract of include/indus/presolve.hpp).
#include "indus/presolve.hpp"

#include <algorithm>
#include <cmath>
#include <limits>

#include "indus/tolerances.hpp"

namespace indus::presolve {

namespace {

constexpr double kFeasTol = 1e-9;
bool is_neg_inf(double v) { return v <= -1e20; }
bool is_pos_inf(double v) { return v >= 1e20; }

struct Work {
    const Model& M;
    la::SparseMatrixCSR R;                // row view of A
    std::vector<char> row_on, col_on;
    std::vector<int> row_cnt, col_cnt;    // active nonzeros per row / column
    std::vector<double> rl, ru, cl, cu;
    double offset;
    explicit Work(const Model& m)
        : M(m), R(m.A.to_csr()), row_on(m.num_rows, 1), col_on(m.num_cols, 1), row_cnt(m.num_rows, 0),
          col_cnt(m.num_cols, 0), rl(m.row_lower), ru(m.row_upper), cl(m.col_lower), cu(m.col_upper),
          offset(m.objective_offset) {
        for (int j = 0; j < m.num_cols; ++j) col_cnt[j] = static_cast<int>(m.A.col_ptr[j + 1] - m.A.col_ptr[j]);
        for (int i = 0; i < m.num_rows; ++i) row_cnt[i] = static_cast<int>(R.row_ptr[i + 1] - R.row_ptr[i]);
    }
    double canon_c(int j) const { return M.sense == ObjSense::kMaximize ? -M.c[j] : M.c[j]; }
    void drop_col(int j) {
        col_on[j] = 0;
        for (int64_t k = M.A.col_ptr[j]; k < M.A.col_ptr[j + 1]; ++k)
            if (row_on[M.A.row_idx[k]]) --row_cnt[M.A.row_idx[k]];
    }
    void drop_row(int i) {
        row_on[i] = 0;
        for (int64_t k = R.row_ptr[i]; k < R.row_ptr[i + 1]; ++k)
            if (col_on[R.col_idx[k]]) --col_cnt[R.col_idx[k]];
    }
};

}  // namespace

PresolveResult PresolveEngine::apply(const Model& model, int max_passes) {
    PresolveResult res;
    Work w(model);
    const int m = model.num_rows, n = model.num_cols;
    auto fail = [&](bool infeasible, const std::string& msg) {
        res.is_infeasible = infeasible;
        res.is_unbounded = !infeasible;
        res.status_message = msg;
        return res;
    };
    for (int pass = 0; pass < max_passes; ++pass) {
        bool changed = false;
        for (int i = 0; i < m; ++i) {  // empty rows
            if (!w.row_on[i] || w.row_cnt[i] != 0) continue;
            if (w.rl[i] > kFeasTol || w.ru[i] < -kFeasTol) return fail(true, "empty row " + std::to_string(i) + " cannot hold 0");
            ReductionRecord r;
            r.type = ReductionType::kEmptyRow;
            r.row_idx = i;
            res.stack.push(r);
            w.drop_row(i);
            ++res.num_empty_rows;
            changed = true;
        }
        for (int j = 0; j < n; ++j) {  // fixed columns, then empty columns
            if (!w.col_on[j]) continue;
            const bool fixed = std::fabs(w.cu[j] - w.cl[j]) <= tol::kZeroDrop && !is_neg_inf(w.cl[j]);
            if (!fixed && w.col_cnt[j] != 0) continue;
            double x;
            ReductionRecord r;
            r.col_idx = j;
            r.obj_cost = model.c[j];
            if (fixed) {
                x = w.cl[j];
                r.type = ReductionType::kFixedCol;
                for (int64_t k = model.A.col_ptr[j]; k < model.A.col_ptr[j + 1]; ++k) {
                    const int i = model.A.row_idx[k];
                    if (!w.row_on[i]) continue;
                    const double a = model.A.values[k];
                    if (!is_neg_inf(w.rl[i])) w.rl[i] -= a * x;
                    if (!is_pos_inf(w.ru[i])) w.ru[i] -= a * x;
                    r.entries.emplace_back(i, a);
                }
                ++res.num_fixed_cols;
            } else {
                const double cj = w.canon_c(j);
                if (cj > tol::kDualFeasibility) {
                    if (is_neg_inf(w.cl[j])) return fail(false, "empty column " + std::to_string(j) + " unbounded below");
                    x = w.cl[j];
                } else if (cj < -tol::kDualFeasibility) {
                    if (is_pos_inf(w.cu[j])) return fail(false, "empty column " + std::to_string(j) + " unbounded above");
                    x = w.cu[j];
                } else {
                    x = std::clamp(0.0, w.cl[j], w.cu[j]);
                }
                r.type = ReductionType::kEmptyCol;
                ++res.num_empty_cols;
            }
            r.fixed_value = x;
            w.offset += model.c[j] * x;
            res.stack.push(r);
            w.drop_col(j);
            changed = true;
        }
        for (int i = 0; i < m; ++i) {  // singleton rows -> column bounds
            if (!w.row_on[i] || w.row_cnt[i] != 1) continue;
            int j = -1;
            double a = 0.0;
            for (int64_t k = w.R.row_ptr[i]; k < w.R.row_ptr[i + 1]; ++k)
                if (w.col_on[w.R.col_idx[k]]) {
                    j = w.R.col_idx[k];
                    a = w.R.values[k];
                }
            if (j < 0 || a == 0.0) continue;
            const double lo_b = a > 0 ? w.rl[i] : w.ru[i], hi_b = a > 0 ? w.ru[i] : w.rl[i];
            const double lo = (a > 0 ? is_neg_inf(lo_b) : is_pos_inf(lo_b)) ? -1e20 : lo_b / a;
            const double hi = (a > 0 ? is_pos_inf(hi_b) : is_neg_inf(hi_b)) ? 1e20 : hi_b / a;
            ReductionRecord r;
            r.type = ReductionType::kSingletonRow;
            r.row_idx = i;
            r.col_idx = j;
            r.coeff = a;
            r.old_lower = w.cl[j];
            r.old_upper = w.cu[j];
            r.row_lower = lo;   // implied column bounds
            r.row_upper = hi;
            w.cl[j] = std::max(w.cl[j], lo);
            w.cu[j] = std::min(w.cu[j], hi);
            if (w.cl[j] > w.cu[j] + kFeasTol * (1.0 + std::fabs(w.cl[j])))
                return fail(true, "singleton row " + std::to_string(i) + " conflicts with the bounds of column " + std::to_string(j));
            if (w.cl[j] > w.cu[j]) w.cu[j] = w.cl[j];
            res.stack.push(r);
            w.drop_row(i);
            ++res.num_singleton_rows;
            changed = true;
        }
        for (int i = 0; i < m; ++i) {  // redundant rows
            if (!w.row_on[i]) continue;
            double lo = 0.0, hi = 0.0;
            bool finite = true;
            for (int64_t k = w.R.row_ptr[i]; k < w.R.row_ptr[i + 1] && finite; ++k) {
                const int j = w.R.col_idx[k];
                if (!w.col_on[j]) continue;
                const double a = w.R.values[k];
                const double p = a * (a > 0 ? w.cl[j] : w.cu[j]), q = a * (a > 0 ? w.cu[j] : w.cl[j]);
                if (is_neg_inf(w.cl[j]) || is_pos_inf(w.cu[j])) finite = false;
                lo += p;
                hi += q;
            }
            if (!finite) continue;
            const bool lo_ok = is_neg_inf(w.rl[i]) || w.rl[i] <= lo - kFeasTol * (1.0 + std::fabs(lo));
            const bool hi_ok = is_pos_inf(w.ru[i]) || w.ru[i] >= hi + kFeasTol * (1.0 + std::fabs(hi));
            if (!lo_ok || !hi_ok) continue;
            ReductionRecord r;
            r.type = ReductionType::kRedundantRow;
            r.row_idx = i;
            res.stack.push(r);
            w.drop_row(i);
            ++res.num_redundant_rows;
            changed = true;
        }
        if (!changed) break;
    }

    // reduced model
    PresolveStack& S = res.stack;
    S.orig_to_reduced_row.assign(m, -1);
    S.orig_to_reduced_col.assign(n, -1);
    for (int i = 0; i < m; ++i)
        if (w.row_on[i]) {
            S.orig_to_reduced_row[i] = static_cast<int>(S.reduced_to_orig_row.size());
            S.reduced_to_orig_row.push_back(i);
        }
    for (int j = 0; j < n; ++j)
        if (w.col_on[j]) {
            S.orig_to_reduced_col[j] = static_cast<int>(S.reduced_to_orig_col.size());
            S.reduced_to_orig_col.push_back(j);
        }
    Model& P = res.presolved_model;
    P.name = model.name + "_presolved";
    P.sense = model.sense;
    P.objective_offset = w.offset;
    P.num_rows = static_cast<int>(S.reduced_to_orig_row.size());
    P.num_cols = static_cast<int>(S.reduced_to_orig_col.size());
    std::vector<la::Triplet> trips;
    for (int jr = 0; jr < P.num_cols; ++jr) {
        const int j = S.reduced_to_orig_col[jr];
        P.c.push_back(model.c[j]);
        P.col_lower.push_back(w.cl[j]);
        P.col_upper.push_back(w.cu[j]);
        for (int64_t k = model.A.col_ptr[j]; k < model.A.col_ptr[j + 1]; ++k) {
            const int ir = S.orig_to_reduced_row[model.A.row_idx[k]];
            if (ir >= 0) trips.push_back({ir, jr, model.A.values[k]});
        }
    }
    for (int ir = 0; ir < P.num_rows; ++ir) {
        P.row_lower.push_back(w.rl[S.reduced_to_orig_row[ir]]);
        P.row_upper.push_back(w.ru[S.reduced_to_orig_row[ir]]);
    }
    P.A.set_from_triplets(P.num_rows, P.num_cols, trips);
    return res;
}

Solution PresolveEngine::postsolve(const Solution& red, const PresolveStack& S, const Model& M) {
    const int m = M.num_rows, n = M.num_cols;
    Solution sol = red;
    sol.col_value.assign(n, 0.0);
    sol.row_dual.assign(m, 0.0);
    sol.col_basis_status.assign(n, BasisStatus::kUnknown);
    sol.row_basis_status.assign(m, BasisStatus::kUnknown);
    for (size_t jr = 0; jr < S.reduced_to_orig_col.size() && jr < red.col_value.size(); ++jr)
        sol.col_value[S.reduced_to_orig_col[jr]] = red.col_value[jr];
    for (size_t ir = 0; ir < S.reduced_to_orig_row.size() && ir < red.row_dual.size(); ++ir)
        sol.row_dual[S.reduced_to_orig_row[ir]] = red.row_dual[ir];
    const auto& recs = S.records();
    for (const auto& r : recs)
        if (r.type == ReductionType::kFixedCol || r.type == ReductionType::kEmptyCol) sol.col_value[r.col_idx] = r.fixed_value;
    // undo in LIFO order: a singleton row takes over the reduced cost of its column when
    // the column sits on the bound the row implied
    const bool is_max = M.sense == ObjSense::kMaximize;
    for (auto it = recs.rbegin(); it != recs.rend(); ++it) {
        if (it->type != ReductionType::kSingletonRow) continue;
        const int i = it->row_idx, j = it->col_idx;
        double d = is_max ? -M.c[j] : M.c[j];
        for (int64_t k = M.A.col_ptr[j]; k < M.A.col_ptr[j + 1]; ++k)
            if (M.A.row_idx[k] != i) d -= M.A.values[k] * sol.row_dual[M.A.row_idx[k]];
        const double xj = sol.col_value[j];
        const double tol = 1e-7 * (1.0 + std::fabs(xj));
        const bool at_implied_lo = it->row_lower > -1e20 && std::fabs(xj - it->row_lower) <= tol && it->row_lower >= it->old_lower - tol;
        const bool at_implied_hi = it->row_upper < 1e20 && std::fabs(xj - it->row_upper) <= tol && it->row_upper <= it->old_upper + tol;
        if ((at_implied_lo && d > 0.0) || (at_implied_hi && d < 0.0)) sol.row_dual[i] = d / it->coeff;
    }
    double obj = M.objective_offset;
    for (int j = 0; j < n; ++j) obj += M.c[j] * sol.col_value[j];
    sol.objective_value = obj;
    sol.best_dual_bound = obj;
    sol.row_value.assign(m, 0.0);
    sol.recompute_quality(M);
    return sol;
}

}  // namespace indus::presolve
