// Presolve and postsolve for the PDHG path (contract of include/indus/presolve.hpp).
//
// The reductions, their order inside a pass, their tests and tolerances, and the postsolve
// formulas follow PresolveEngine of the reference (src/presolve/presolve.cpp:10-976), so the
// reduced model and the reconstructed solution are those the reference produces. The working
// matrix is held as per-row and per-column hash maps (index -> coefficient) that are updated
// in place by the fixed-column, forcing-row, free-column-singleton and doubleton-row
// reductions; hash maps are also what fixes the order in which a row's or a column's entries
// are visited (the order decides, for instance, which two columns a doubleton row names
// first), so they are kept here with the same sequence of insertions and erasures.
#include "indus/presolve.hpp"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <memory>
#include <string>
#include <unordered_map>

#include "indus/tolerances.hpp"

namespace indus::presolve {

namespace {

// Bump allocator for the hash maps of one presolve call: nodes and bucket arrays come from
// large blocks and are released together when the call ends (an erased node is not reused).
// Only where memory comes from changes; the maps' visiting order does not depend on it.
class Arena {
public:
    void* get(size_t bytes, size_t align) {
        size_t pad = (align - reinterpret_cast<uintptr_t>(cur_) % align) % align;
        if (pad + bytes > left_) {
            const size_t sz = std::max<size_t>(bytes + align, size_t(4) << 20);
            blocks_.emplace_back(new char[sz]);
            cur_ = blocks_.back().get();
            left_ = sz;
            pad = (align - reinterpret_cast<uintptr_t>(cur_) % align) % align;
        }
        char* p = cur_ + pad;
        cur_ = p + bytes;
        left_ -= pad + bytes;
        return p;
    }

private:
    std::vector<std::unique_ptr<char[]>> blocks_;
    char* cur_ = nullptr;
    size_t left_ = 0;
};

template <class T>
struct ArenaAlloc {
    using value_type = T;
    Arena* arena;
    explicit ArenaAlloc(Arena* a) : arena(a) {}
    template <class U>
    ArenaAlloc(const ArenaAlloc<U>& o) : arena(o.arena) {}
    T* allocate(size_t k) { return static_cast<T*>(arena->get(k * sizeof(T), alignof(T))); }
    void deallocate(T*, size_t) noexcept {}
    template <class U>
    bool operator==(const ArenaAlloc<U>& o) const { return arena == o.arena; }
    template <class U>
    bool operator!=(const ArenaAlloc<U>& o) const { return arena != o.arena; }
};

using Entries = std::unordered_map<int, double, std::hash<int>, std::equal_to<int>,
                                   ArenaAlloc<std::pair<const int, double>>>;
constexpr double kInf = 1e20;
constexpr double kInfTest = 1e19;   // bounds at or beyond +-1e19 count as infinite

class Reducer {
public:
    Reducer(const Model& m, PresolveResult& r)
        : M(m), res(r), m_(m.num_rows), n_(m.num_cols), row_on(m_, true), col_on(n_, true),
          row_touched(m_, false), rl(m.row_lower), ru(m.row_upper), cl(m.col_lower), cu(m.col_upper),
          cost(m.c), offset(m.objective_offset), row_dirty(m_, 1) {
        // each map built empty with the allocator alone, as a default-constructed map
        const ArenaAlloc<std::pair<const int, double>> alloc(&arena);
        rows.reserve(m_);
        cols.reserve(n_);
        for (int i = 0; i < m_; ++i) rows.emplace_back(alloc);
        for (int j = 0; j < n_; ++j) cols.emplace_back(alloc);
    }

    // false: presolve proved infeasibility or unboundedness (res carries the message)
    bool run(int max_passes) {
        if (!bounds_consistent()) return false;
        // Each row map receives its columns in ascending order and each column map its rows
        // in the stored order, the insertion sequences of one column sweep (presolve.cpp:
        // 65-77); filling map after map keeps each map's nodes together in the arena.
        for (int j = 0; j < n_; ++j) {
            const auto rr = M.A.col_rows(j);
            const auto vv = M.A.col_vals(j);
            for (size_t k = 0; k < rr.size(); ++k)
                if (std::fabs(vv[k]) > tol::kZeroDrop) cols[j][rr[k]] = vv[k];
        }
        const la::SparseMatrixCSR R = M.A.to_csr();
        for (int i = 0; i < m_; ++i)
            for (int64_t k = R.row_ptr[i]; k < R.row_ptr[i + 1]; ++k)
                if (std::fabs(R.values[k]) > tol::kZeroDrop) rows[i][R.col_idx[k]] = R.values[k];
        for (int pass = 0; pass < max_passes; ++pass) {
            int done = 0;
            if (!empty_rows(done) || !empty_cols(done) || !fixed_cols(done) || !singleton_rows(done) ||
                !activity_rows(done) || !free_col_singletons(done) || !doubleton_rows(done))
                return false;
            if (done == 0) break;
        }
        build_reduced();
        return true;
    }

private:
    const Model& M;
    PresolveResult& res;
    int m_, n_;
    std::vector<bool> row_on, col_on, row_touched;
    std::vector<double> rl, ru, cl, cu, cost;
    double offset;
    // Rows whose activity test could decide differently than at their last evaluation: a
    // row's entries, its bounds or a bound of one of its columns changed since. A clean row
    // is skipped by the activity step, which would reach the same "no reduction" again.
    std::vector<char> row_dirty;
    Arena arena;   // declared before the maps: they are destroyed first
    std::vector<Entries> rows, cols;

    void dirty_rows_of(int j) {
        for (const auto& e : cols[j]) row_dirty[e.first] = 1;
    }

    std::string rname(int i) const {
        if (static_cast<size_t>(i) < M.row_names.size() && !M.row_names[i].empty()) return M.row_names[i];
        return "r" + std::to_string(i);
    }
    std::string cname(int j) const {
        if (static_cast<size_t>(j) < M.col_names.size() && !M.col_names[j].empty()) return M.col_names[j];
        return "c" + std::to_string(j);
    }
    bool infeasible(std::string msg) {
        res.is_infeasible = true;
        res.status_message = std::move(msg);
        return false;
    }
    bool unbounded(std::string msg) {
        res.is_unbounded = true;
        res.status_message = std::move(msg);
        return false;
    }
    // remove row i from the column maps of its entries, then empty it
    void detach_row(int i) {
        for (const auto& e : rows[i]) cols[e.first].erase(i);
        rows[i].clear();
    }

    bool bounds_consistent() {   // presolve.cpp:31-45
        for (int j = 0; j < n_; ++j)
            if (M.col_lower[j] > M.col_upper[j] + tol::kZeroDrop) return infeasible("Crossed bounds on column " + cname(j));
        for (int i = 0; i < m_; ++i)
            if (M.row_lower[i] > M.row_upper[i] + tol::kZeroDrop) return infeasible("Crossed bounds on row " + rname(i));
        return true;
    }

    bool empty_rows(int& done) {   // presolve.cpp:83-107
        for (int i = 0; i < m_; ++i) {
            if (!row_on[i] || !rows[i].empty()) continue;
            if (rl[i] > tol::kPrimalFeasibility || ru[i] < -tol::kPrimalFeasibility)
                return infeasible("Empty row " + rname(i) + " cannot be satisfied: [ " + std::to_string(rl[i]) + ", " +
                                  std::to_string(ru[i]) + " ] does not contain 0");
            row_on[i] = false;
            ReductionRecord r;
            r.type = ReductionType::kEmptyRow;
            r.row_idx = i;
            r.name = rname(i);
            r.row_lower = rl[i];
            r.row_upper = ru[i];
            res.stack.push(std::move(r));
            ++res.num_empty_rows;
            ++done;
        }
        return true;
    }

    // An empty column goes to the bound its cost favours, as the raw cost (not the
    // sense-adjusted one) decides: presolve.cpp:109-148.
    bool empty_cols(int& done) {
        for (int j = 0; j < n_; ++j) {
            if (!col_on[j] || !cols[j].empty()) continue;
            col_on[j] = false;
            const double cj = cost[j], lj = cl[j], uj = cu[j];
            double v;
            if (cj > tol::kDualFeasibility) {
                if (lj <= -kInfTest)
                    return unbounded("Unbounded: empty column " + cname(j) + " with positive cost has no lower bound");
                v = lj;
            } else if (cj < -tol::kDualFeasibility) {
                if (uj >= kInfTest)
                    return unbounded("Unbounded: empty column " + cname(j) + " with negative cost has no upper bound");
                v = uj;
            } else {
                v = lj > -kInfTest ? lj : (uj < kInfTest ? uj : 0.0);
            }
            offset += cj * v;
            ReductionRecord r;
            r.type = ReductionType::kEmptyCol;
            r.col_idx = j;
            r.name = cname(j);
            r.fixed_value = v;
            r.obj_cost = cj;
            res.stack.push(std::move(r));
            ++res.num_empty_cols;
            ++done;
        }
        return true;
    }

    bool fixed_cols(int& done) {   // presolve.cpp:150-189
        for (int j = 0; j < n_; ++j) {
            if (!col_on[j] || std::fabs(cu[j] - cl[j]) > tol::kZeroDrop) continue;
            col_on[j] = false;
            const double v = cl[j];
            offset += cost[j] * v;
            ReductionRecord r;
            r.type = ReductionType::kFixedCol;
            r.col_idx = j;
            r.name = cname(j);
            r.fixed_value = v;
            r.obj_cost = cost[j];
            for (const auto& e : cols[j]) {
                const int i = e.first;
                const double a = e.second;
                r.entries.emplace_back(i, a);
                if (rl[i] > -kInfTest) rl[i] -= a * v;
                if (ru[i] < kInfTest) ru[i] -= a * v;
                rows[i].erase(j);
                row_dirty[i] = 1;
                if (rl[i] > ru[i] + tol::kZeroDrop)
                    return infeasible("Fixed variable " + cname(j) + " created crossed bounds on row " + rname(i));
            }
            cols[j].clear();
            res.stack.push(std::move(r));
            ++res.num_fixed_cols;
            ++done;
        }
        return true;
    }

    bool singleton_rows(int& done) {   // presolve.cpp:191-245
        for (int i = 0; i < m_; ++i) {
            if (!row_on[i] || rows[i].size() != 1) continue;
            const int j = rows[i].begin()->first;
            const double a = rows[i].begin()->second;
            if (std::fabs(a) <= tol::kZeroDrop) continue;
            const double li = rl[i], ui = ru[i];
            // implied bounds of x_j from li <= a x_j <= ui: the side that bounds x_j from
            // below is li for a > 0 and ui for a < 0
            const double below = a > 0.0 ? li : ui, above = a > 0.0 ? ui : li;
            const double impl_lo = (a > 0.0 ? li > -kInfTest : ui < kInfTest) ? below / a : -kInf;
            const double impl_hi = (a > 0.0 ? ui < kInfTest : li > -kInfTest) ? above / a : kInf;
            const double old_lo = cl[j], old_hi = cu[j];
            const double new_lo = std::max(old_lo, impl_lo), new_hi = std::min(old_hi, impl_hi);
            if (new_lo > new_hi + tol::kZeroDrop)
                return infeasible("Singleton row " + rname(i) + " creates crossed bounds on column " + cname(j));
            cl[j] = new_lo;
            cu[j] = new_hi;
            row_on[i] = false;
            rows[i].clear();
            cols[j].erase(i);
            dirty_rows_of(j);
            ReductionRecord r;
            r.type = ReductionType::kSingletonRow;
            r.row_idx = i;
            r.col_idx = j;
            r.coeff = a;
            r.row_lower = li;
            r.row_upper = ui;
            r.old_lower = old_lo;
            r.old_upper = old_hi;
            r.name = rname(i);
            res.stack.push(std::move(r));
            ++res.num_singleton_rows;
            ++done;
        }
        return true;
    }

    // Activity bounds of each row with two or more entries: infeasibility, redundant rows
    // (both sides strictly looser than the activity range, 1e-5 margin, not equalities) and
    // forcing rows (one side equal to the extreme activity: every column goes to the bound
    // that attains it). presolve.cpp:247-346.
    bool activity_rows(int& done) {
        for (int i = 0; i < m_; ++i) {
            if (!row_on[i] || rows[i].size() <= 1 || !row_dirty[i]) continue;
            row_dirty[i] = 0;
            double amin = 0.0, amax = 0.0;
            bool min_inf = false, max_inf = false;
            for (const auto& e : rows[i]) {
                const double a = e.second, lj = cl[e.first], uj = cu[e.first];
                const double at_min = a > 0.0 ? lj : uj, at_max = a > 0.0 ? uj : lj;
                const bool min_unb = a > 0.0 ? lj <= -kInfTest : uj >= kInfTest;
                const bool max_unb = a > 0.0 ? uj >= kInfTest : lj <= -kInfTest;
                if (min_unb) min_inf = true;
                else amin += a * at_min;
                if (max_unb) max_inf = true;
                else amax += a * at_max;
            }
            const double li = rl[i], ui = ru[i];
            if (!min_inf && amin > ui + tol::kPrimalFeasibility)
                return infeasible("Minimum activity of row " + rname(i) + " exceeds upper bound");
            if (!max_inf && amax < li - tol::kPrimalFeasibility)
                return infeasible("Maximum activity of row " + rname(i) + " is below lower bound");
            const bool equality = std::fabs(li - ui) <= tol::kZeroDrop;
            const bool upper_loose = ui >= kInfTest || (!max_inf && amax < ui - 1e-5);
            const bool lower_loose = li <= -kInfTest || (!min_inf && amin > li + 1e-5);
            if (!equality && upper_loose && lower_loose) {
                row_on[i] = false;
                detach_row(i);
                ReductionRecord r;
                r.type = ReductionType::kRedundantRow;
                r.row_idx = i;
                r.name = rname(i);
                res.stack.push(std::move(r));
                ++res.num_redundant_rows;
                ++done;
                continue;
            }
            const bool at_lower = !max_inf && std::fabs(amax - li) <= tol::kPrimalFeasibility;
            const bool at_upper = !min_inf && std::fabs(amin - ui) <= tol::kPrimalFeasibility;
            if (!at_lower && !at_upper) continue;
            ReductionRecord r;
            r.type = ReductionType::kForcingRow;
            r.row_idx = i;
            r.name = rname(i);
            r.row_lower = li;
            r.row_upper = ui;
            std::vector<std::pair<int, double>> fix;
            for (const auto& e : rows[i]) {
                const int j = e.first;
                const double a = e.second;
                r.entries.emplace_back(j, a);
                // at_lower: the row needs its maximum activity; at_upper: its minimum
                fix.emplace_back(j, (a > 0.0) == at_lower ? cu[j] : cl[j]);
            }
            row_on[i] = false;
            detach_row(i);
            res.stack.push(std::move(r));
            ++res.num_forcing_rows;
            ++done;
            for (const auto& f : fix) {
                cl[f.first] = cu[f.first] = f.second;
                dirty_rows_of(f.first);
            }
        }
        return true;
    }

    // A free column met only by one equality row: the row defines it, and its cost moves
    // onto the row's other columns. presolve.cpp:348-391.
    bool free_col_singletons(int& done) {
        for (int j = 0; j < n_; ++j) {
            if (!col_on[j] || cl[j] > -kInfTest || cu[j] < kInfTest || cols[j].size() != 1) continue;
            const int i = cols[j].begin()->first;
            const double a = cols[j].begin()->second;
            if (!row_on[i] || std::fabs(rl[i] - ru[i]) > tol::kPrimalFeasibility || std::fabs(a) <= 1e-4) continue;
            const double b = rl[i], cj = cost[j];
            ReductionRecord r;
            r.type = ReductionType::kFreeColSingleton;
            r.row_idx = i;
            r.col_idx = j;
            r.coeff = a;
            r.fixed_value = b;
            r.obj_cost = cj;
            r.name = cname(j);
            offset += cj * (b / a);
            for (const auto& e : rows[i]) {
                if (e.first == j) continue;
                r.entries.emplace_back(e.first, e.second);
                cost[e.first] -= cj * (e.second / a);
                cols[e.first].erase(i);
            }
            col_on[j] = false;
            row_on[i] = false;
            cols[j].clear();
            rows[i].clear();
            res.stack.push(std::move(r));
            ++res.num_free_col_singletons;
            ++done;
        }
        return true;
    }

    // a column whose every original entry is still in place (all its rows active)
    bool untouched_col(int j) const {
        const auto rr = M.A.col_rows(j);
        if (cols[j].size() != rr.size()) return false;
        for (size_t k = 0; k < rr.size(); ++k)
            if (!row_on[rr[k]]) return false;
        return true;
    }

    // Equality rows a1 x1 + a2 x2 = b: one column is eliminated through the other and
    // substituted into its (at most one) other row. presolve.cpp:393-566.
    bool doubleton_rows(int& done) {
        for (int i = 0; i < m_; ++i) {
            if (!row_on[i] || row_touched[i] || rows[i].size() != 2 || std::fabs(rl[i] - ru[i]) > tol::kZeroDrop) continue;
            auto it = rows[i].begin();
            const int j1 = it->first;
            const double a1 = it->second;
            ++it;
            const int j2 = it->first;
            const double a2 = it->second;
            const double b = rl[i];
            if (!std::isfinite(a1) || !std::isfinite(a2) || !std::isfinite(b)) continue;
            if (std::fabs(a1) < 1e-6 || std::fabs(a2) < 1e-6) continue;
            if (std::max(std::fabs(a1), std::fabs(a2)) / std::min(std::fabs(a1), std::fabs(a2)) > 1e4) continue;
            const bool u1 = untouched_col(j1), u2 = untouched_col(j2);
            if (!u1 && !u2) continue;
            // prefer the untouched column, then the one in fewer rows, then the larger |a|
            bool first = u1;
            if (u1 && u2) {
                const size_t d1 = cols[j1].size(), d2 = cols[j2].size();
                first = d1 != d2 ? d1 < d2 : std::fabs(a1) >= std::fabs(a2);
            }
            const int je = first ? j1 : j2, jk = first ? j2 : j1;
            const double ae = first ? a1 : a2, ak = first ? a2 : a1;
            if (cols[je].size() > 2 || std::fabs(ae) < 1e-4) continue;
            // x_keep = (b - ae x_elim) / ak carries x_elim's bounds over to x_keep
            const double le = cl[je], ue = cu[je];
            const bool rising = -ae / ak > 0.0;
            double impl_lo = -kInf, impl_hi = kInf;
            const double lo_src = rising ? le : ue, hi_src = rising ? ue : le;
            const bool lo_fin = rising ? le > -kInfTest : ue < kInfTest;
            const bool hi_fin = rising ? ue < kInfTest : le > -kInfTest;
            if (lo_fin) impl_lo = (b - ae * lo_src) / ak;
            if (hi_fin) impl_hi = (b - ae * hi_src) / ak;
            const double nlo = std::max(cl[jk], impl_lo), nhi = std::min(cu[jk], impl_hi);
            if (nlo > nhi + tol::kZeroDrop)
                return infeasible("Doubleton row " + rname(i) + " creates crossed bounds on " + cname(jk));
            cl[jk] = nlo;
            cu[jk] = nhi;
            ReductionRecord r;
            r.type = ReductionType::kDoubletonRow;
            r.row_idx = i;
            r.col_idx = je;
            r.other_col_idx = jk;
            r.coeff = ae;
            r.other_coeff = ak;
            r.fixed_value = b;
            r.obj_cost = cost[je];
            r.old_lower = le;
            r.old_upper = ue;
            r.name = rname(i);
            const double ce = cost[je];
            offset += ce * (b / ae);
            cost[jk] -= ce * (ak / ae);
            cost[je] = 0.0;
            std::vector<std::pair<int, double>> others;
            for (const auto& e : cols[je])
                if (e.first != i) {
                    others.emplace_back(e.first, e.second);
                    r.entries.emplace_back(e.first, e.second);
                }
            std::string crossed;
            for (const auto& o : others) {
                const int q = o.first;
                const double aq = o.second;
                row_touched[q] = true;
                row_dirty[q] = 1;
                const double shift = aq * (b / ae);
                if (rl[q] > -kInfTest) rl[q] -= shift;
                if (ru[q] < kInfTest) ru[q] -= shift;
                if (rl[q] > ru[q] + tol::kZeroDrop) crossed = rname(q);
                double& merged = rows[q][jk];   // creates the entry when x_keep was absent
                merged += -aq * (ak / ae);
                if (std::fabs(merged) <= tol::kZeroDrop) {
                    rows[q].erase(jk);
                    cols[jk].erase(q);
                } else {
                    cols[jk][q] = rows[q][jk];
                }
                rows[q].erase(je);
            }
            if (!crossed.empty()) return infeasible("Doubleton substitution created crossed bounds on row " + crossed);
            col_on[je] = false;
            row_on[i] = false;
            cols[je].clear();
            rows[i].clear();
            cols[jk].erase(i);
            dirty_rows_of(jk);   // x_keep's bounds were tightened
            res.stack.push(std::move(r));
            ++res.num_doubleton_rows;
            ++done;
        }
        return true;
    }

    void build_reduced() {   // presolve.cpp:574-645
        PresolveStack& S = res.stack;
        S.orig_to_reduced_row.assign(m_, -1);
        S.orig_to_reduced_col.assign(n_, -1);
        for (int i = 0; i < m_; ++i)
            if (row_on[i]) {
                S.orig_to_reduced_row[i] = static_cast<int>(S.reduced_to_orig_row.size());
                S.reduced_to_orig_row.push_back(i);
            }
        for (int j = 0; j < n_; ++j)
            if (col_on[j]) {
                S.orig_to_reduced_col[j] = static_cast<int>(S.reduced_to_orig_col.size());
                S.reduced_to_orig_col.push_back(j);
            }
        Model P;
        P.name = M.name + "_presolved";
        P.sense = M.sense;
        P.objective_offset = offset;
        P.num_rows = static_cast<int>(S.reduced_to_orig_row.size());
        P.num_cols = static_cast<int>(S.reduced_to_orig_col.size());
        // the reduced matrix, column by column with rows ascending: what assembling the
        // surviving entries as triplets gives (no duplicates occur, |a| <= 1e-11 dropped)
        P.A.resize(P.num_rows, P.num_cols);
        std::vector<std::pair<int, double>> colbuf;
        for (int jr = 0; jr < P.num_cols; ++jr) {
            const int j = S.reduced_to_orig_col[jr];
            P.c.push_back(cost[j]);
            P.col_lower.push_back(cl[j]);
            P.col_upper.push_back(cu[j]);
            P.col_names.push_back(cname(j));
            colbuf.clear();
            for (const auto& e : cols[j]) {
                const int ir = S.orig_to_reduced_row[e.first];
                if (ir >= 0 && std::fabs(e.second) > tol::kZeroDrop) colbuf.emplace_back(ir, e.second);
            }
            std::sort(colbuf.begin(), colbuf.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
            P.A.col_ptr[jr] = static_cast<int64_t>(P.A.row_idx.size());
            for (const auto& e : colbuf) {
                P.A.row_idx.push_back(e.first);
                P.A.values.push_back(e.second);
            }
        }
        P.A.col_ptr[P.num_cols] = static_cast<int64_t>(P.A.row_idx.size());
        for (int ir = 0; ir < P.num_rows; ++ir) {
            const int i = S.reduced_to_orig_row[ir];
            P.row_lower.push_back(rl[i]);
            P.row_upper.push_back(ru[i]);
            P.row_names.push_back(rname(i));
        }
        res.presolved_model = std::move(P);
    }
};

// sum over the original column j of a_rj y_r, leaving out row `skip`
double col_dot_except(const Model& M, int j, int skip, const std::vector<double>& y) {
    const auto rr = M.A.col_rows(j);
    const auto vv = M.A.col_vals(j);
    double s = 0.0;
    for (size_t k = 0; k < rr.size(); ++k)
        if (rr[k] != skip) s += vv[k] * y[rr[k]];
    return s;
}

}  // namespace

PresolveResult PresolveEngine::apply(const Model& model, int max_passes) {
    PresolveResult res;
    res.presolved_model = model;   // what an infeasible / unbounded result carries
    Reducer red(model, res);
    red.run(max_passes);
    return res;
}

// Rebuild the original solution: retained entries mapped back, then the reductions undone
// newest first (presolve.cpp:649-959), then A x, the reduced costs and the objective
// recomputed on the original model.
Solution PresolveEngine::postsolve(const Solution& red, const PresolveStack& S, const Model& M) {
    const int m = M.num_rows, n = M.num_cols;
    const double sense = M.sense == ObjSense::kMaximize ? -1.0 : 1.0;
    Solution s;
    s.status = red.status;
    s.status_message = red.status_message;
    s.iterations = red.iterations;
    s.nodes = red.nodes;
    s.algorithm_used = red.algorithm_used;
    s.certificate_type = red.certificate_type;
    s.certificate_vector = red.certificate_vector;
    s.col_value.assign(n, 0.0);
    s.col_dual.assign(n, 0.0);
    s.row_value.assign(m, 0.0);
    s.row_dual.assign(m, 0.0);
    s.col_basis_status.assign(n, BasisStatus::kAtLower);
    s.row_basis_status.assign(m, BasisStatus::kBasic);
    for (size_t jr = 0; jr < S.reduced_to_orig_col.size(); ++jr) {
        const int j = S.reduced_to_orig_col[jr];
        if (jr < red.col_value.size()) s.col_value[j] = red.col_value[jr];
        if (jr < red.col_dual.size()) s.col_dual[j] = red.col_dual[jr];
        if (jr < red.col_basis_status.size()) s.col_basis_status[j] = red.col_basis_status[jr];
    }
    for (size_t ir = 0; ir < S.reduced_to_orig_row.size(); ++ir) {
        const int i = S.reduced_to_orig_row[ir];
        if (ir < red.row_value.size()) s.row_value[i] = red.row_value[ir];
        if (ir < red.row_dual.size()) s.row_dual[i] = red.row_dual[ir];
        if (ir < red.row_basis_status.size()) s.row_basis_status[i] = red.row_basis_status[ir];
    }
    // sign restriction of a row's dual for a one-sided row (canonical minimisation)
    auto one_sided = [](double lo, double hi, double y) {
        if (lo <= -kInfTest && hi < kInfTest) return std::min(y, 0.0);   // <= row
        if (lo > -kInfTest && hi >= kInfTest) return std::max(y, 0.0);   // >= row
        return y;
    };
    const auto& recs = S.records();
    for (auto it = recs.rbegin(); it != recs.rend(); ++it) {
        const ReductionRecord& r = *it;
        switch (r.type) {
            case ReductionType::kFixedCol:
                s.col_value[r.col_idx] = r.fixed_value;
                s.col_basis_status[r.col_idx] = BasisStatus::kFixed;
                break;
            case ReductionType::kEmptyCol:
                s.col_value[r.col_idx] = r.fixed_value;
                s.col_dual[r.col_idx] = r.obj_cost;
                s.col_basis_status[r.col_idx] = BasisStatus::kAtLower;
                break;
            case ReductionType::kEmptyRow:
                s.row_value[r.row_idx] = 0.0;
                s.row_dual[r.row_idx] = 0.0;
                s.row_basis_status[r.row_idx] = BasisStatus::kBasic;
                break;
            case ReductionType::kSingletonRow: {
                // the row prices x_j's reduced cost when x_j rests on a bound the row implied
                const int i = r.row_idx, j = r.col_idx;
                const double a = r.coeff, xj = s.col_value[j];
                s.row_value[i] = a * xj;
                if (std::fabs(a) <= 1e-15) break;
                const double d0 = sense * M.c[j] - col_dot_except(M, j, i, s.row_dual);
                const double lj = M.col_lower[j], uj = M.col_upper[j];
                const bool equality = std::fabs(r.row_lower - r.row_upper) <= tol::kZeroDrop;
                double y = 0.0;
                if (equality) {
                    y = d0 / a;
                    s.row_basis_status[i] = BasisStatus::kFixed;
                } else if ((xj > lj + 1e-5 && xj < uj - 1e-5) || (xj <= lj + 1e-5 && d0 < -tol::kDualFeasibility) ||
                           (xj >= uj - 1e-5 && d0 > tol::kDualFeasibility)) {
                    y = d0 / a;
                }
                if (!equality) y = one_sided(r.row_lower, r.row_upper, y);
                s.row_dual[i] = y;
                break;
            }
            case ReductionType::kForcingRow: {
                // the interval of row duals that keeps every fixed column's reduced cost
                // sign-consistent with the bound it was fixed at; pick the point nearest 0
                const int i = r.row_idx;
                const bool equality = std::fabs(r.row_lower - r.row_upper) <= tol::kZeroDrop;
                double ylo = -kInf, yhi = kInf;
                if (!equality) {
                    if (r.row_lower <= -kInfTest && r.row_upper < kInfTest) yhi = std::min(yhi, 0.0);
                    else if (r.row_lower > -kInfTest && r.row_upper >= kInfTest) ylo = std::max(ylo, 0.0);
                }
                for (const auto& e : r.entries) {
                    const int j = e.first;
                    const double a = e.second;
                    if (std::fabs(a) <= tol::kZeroDrop) continue;
                    const double ct = sense * M.c[j] - col_dot_except(M, j, i, s.row_dual);
                    const double xj = s.col_value[j];
                    const bool at_lo = xj <= M.col_lower[j] + 1e-5, at_hi = xj >= M.col_upper[j] - 1e-5;
                    const double t = ct / a;
                    if (at_lo && !at_hi) {          // d_j >= 0
                        if (a > 0.0) yhi = std::min(yhi, t);
                        else ylo = std::max(ylo, t);
                    } else if (at_hi && !at_lo) {   // d_j <= 0
                        if (a > 0.0) ylo = std::max(ylo, t);
                        else yhi = std::min(yhi, t);
                    } else if (!at_lo && !at_hi) {  // d_j = 0
                        ylo = std::max(ylo, t);
                        yhi = std::min(yhi, t);
                    }
                }
                double y;
                if (ylo <= yhi) y = (ylo <= 0.0 && yhi >= 0.0) ? 0.0 : (ylo > 0.0 ? ylo : yhi);
                else y = 0.5 * (ylo + yhi);
                s.row_dual[i] = y;
                s.row_basis_status[i] = BasisStatus::kFixed;
                break;
            }
            case ReductionType::kRedundantRow:
                s.row_dual[r.row_idx] = 0.0;
                s.row_basis_status[r.row_idx] = BasisStatus::kBasic;
                break;
            case ReductionType::kFreeColSingleton: {
                double other = 0.0;
                for (const auto& e : r.entries) other += e.second * s.col_value[e.first];
                s.col_value[r.col_idx] = (r.fixed_value - other) / r.coeff;
                s.row_dual[r.row_idx] = r.obj_cost / r.coeff;
                s.col_dual[r.col_idx] = 0.0;
                s.col_basis_status[r.col_idx] = BasisStatus::kBasic;
                break;
            }
            case ReductionType::kDoubletonRow: {
                const int je = r.col_idx, jk = r.other_col_idx, i = r.row_idx;
                const double ae = r.coeff, ak = r.other_coeff, b = r.fixed_value;
                const double xk = s.col_value[jk];
                const double xe = (b - ak * xk) / ae;
                s.col_value[je] = xe;
                double ae_y = 0.0;
                for (const auto& e : r.entries) ae_y += e.second * s.row_dual[e.first];
                const double ce_t = sense * r.obj_cost - ae_y;
                const double lk = M.col_lower[jk], uk = M.col_upper[jk];
                const bool k_interior = !(lk > -kInfTest && xk <= lk + 1e-5) && !(uk < kInfTest && xk >= uk - 1e-5);
                const bool e_lo = r.old_lower > -kInfTest && xe <= r.old_lower + 1e-5;
                const bool e_hi = r.old_upper < kInfTest && xe >= r.old_upper - 1e-5;
                const bool e_interior = !e_lo && !e_hi;
                // the kept column's original coefficient in this row, if it had one
                double ak0 = 0.0;
                bool k_in_row = false;
                {
                    const auto rr = M.A.col_rows(jk);
                    const auto vv = M.A.col_vals(jk);
                    for (size_t k = 0; k < rr.size(); ++k)
                        if (rr[k] == i) {
                            k_in_row = true;
                            ak0 = vv[k];
                            break;
                        }
                }
                double y;
                if (k_interior && !e_interior && k_in_row && std::fabs(ak0) > 1e-12)
                    y = (sense * M.c[jk] - col_dot_except(M, jk, i, s.row_dual)) / ak0;   // d_keep = 0
                else
                    y = ce_t / ae;                                                       // d_elim = 0
                s.row_dual[i] = y;
                s.row_basis_status[i] = BasisStatus::kFixed;
                s.col_basis_status[je] = e_interior ? BasisStatus::kBasic
                                                    : (e_lo && e_hi ? BasisStatus::kFixed
                                                                    : (e_lo ? BasisStatus::kAtLower : BasisStatus::kAtUpper));
                break;
            }
        }
    }
    M.A.multiply(s.col_value, s.row_value);
    refine_dual_solution(M, s);
    double obj = M.objective_offset;
    for (int j = 0; j < n; ++j) obj += M.c[j] * s.col_value[j];
    s.objective_value = obj;
    s.best_dual_bound = obj;
    s.recompute_quality(M);
    return s;
}

// Reduced costs from the row duals on the original model: d = sense * c - A^T y, reported
// with the sign of the original objective (presolve.cpp:961-976; tol is unused there too).
void PresolveEngine::refine_dual_solution(const Model& M, Solution& s, double tol) {
    (void)tol;
    std::vector<double> aty(M.num_cols, 0.0);
    M.A.multiply_transpose(s.row_dual, aty);
    const bool is_max = M.sense == ObjSense::kMaximize;
    s.col_dual.resize(M.num_cols);
    for (int j = 0; j < M.num_cols; ++j) {
        const double d = (is_max ? -M.c[j] : M.c[j]) - aty[j];
        s.col_dual[j] = is_max ? -d : d;
    }
}

}  // namespace indus::presolve
