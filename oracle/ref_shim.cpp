// Test infrastructure only: a C-ABI shim over the reference library compiled from
// its own sources (see oracle/Makefile). Used by tests/golden/make_golden.py in the
// build container to write golden vectors, and (optionally) by tests to cross-check
// the C restatement in oracle/pdhg_oracle.c. Never linked into the product.
//
// Reference entry points wrapped:
//   indus::solve                          src/engine/model.cpp:376
//   indus::pdhg::solve_pdhg_cpu           src/solvers/pdhg/pdhg.cpp:448
//   indus::pdhg::estimate_spectral_norm   src/solvers/pdhg/pdhg.cpp:11
//   indus::la::RuizScaler::compute_and_scale  src/linalg/scaling.cpp:7
//   indus::la::SparseMatrixCSC::multiply / multiply_transpose  src/linalg/sparse.cpp:196,219
//   indus::io::read_mps                   src/io/mps_reader.cpp:48
//   indus::presolve::PresolveEngine::apply / postsolve  src/presolve/presolve.cpp:10,649
//   indus::verifier::verify_solution      src/engine/verifier.cpp:25
//   indus::Options (set / get_*)          src/engine/options.cpp:6-89
//   indus::Solution::recompute_quality    src/engine/model.cpp:203-374
//   indus::io::write_solution / write_json  src/io/writer.cpp:85-264
//   indus::verifier::verify_files         src/engine/verifier.cpp:352-668
#include "../tests/cpp/audit_dump.hpp"
#include "../tests/cpp/options_probe.hpp"
#include "../tests/cpp/presolve_dump.hpp"
#include "indus/options.hpp"
#include "indus/presolve.hpp"
#include "indus/model.hpp"
#include "indus/pdhg.hpp"
#include "indus/io.hpp"
#include "indus/verifier.hpp"
#include "src/linalg/scaling.hpp"

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

namespace {

indus::Model make_model(int m, int n, const int64_t* col_ptr, const int* row_idx,
                        const double* vals, const double* c, const double* lc,
                        const double* uc, const double* lr, const double* ur,
                        int maximize, double offset) {
    indus::Model model;
    model.num_rows = m;
    model.num_cols = n;
    model.sense = maximize ? indus::ObjSense::kMaximize : indus::ObjSense::kMinimize;
    model.objective_offset = offset;
    model.c.assign(c, c + n);
    model.col_lower.assign(lc, lc + n);
    model.col_upper.assign(uc, uc + n);
    model.row_lower.assign(lr, lr + m);
    model.row_upper.assign(ur, ur + m);
    // Assemble through the reference's own triplet path so drop rules match.
    std::vector<indus::la::Triplet> trips;
    trips.reserve(static_cast<size_t>(col_ptr[n]));
    for (int j = 0; j < n; ++j)
        for (int64_t k = col_ptr[j]; k < col_ptr[j + 1]; ++k)
            trips.push_back({row_idx[k], j, vals[k]});
    model.A.set_from_triplets(m, n, trips);
    return model;
}

indus::la::SparseMatrixCSC make_csc(int m, int n, const int64_t* col_ptr,
                                    const int* row_idx, const double* vals) {
    indus::la::SparseMatrixCSC A(m, n);
    A.col_ptr.assign(col_ptr, col_ptr + n + 1);
    A.row_idx.assign(row_idx, row_idx + col_ptr[n]);
    A.values.assign(vals, vals + col_ptr[n]);
    return A;
}

}  // namespace

extern "C" {

// scal_out (24 doubles): 0 status, 1 objective, 2 iterations, 3 tau, 4 sigma,
// 5 spectral_norm, 6 restart_count, 7 primal_res, 8 dual_res, 9 gap, 10 converged,
// 11 q.max_primal_violation, 12 q.max_dual_violation, 13 q.is_primal_feasible,
// 14 q.is_dual_feasible, 15 q.max_complementarity_violation, 16 best_dual_bound,
// 17 setup_time, 18 iteration_time, 19 total_time, 20 presolve_total_reductions
// mode: 0 = pdhg::solve_pdhg_cpu directly, 1 = indus::solve (dispatcher)
int ref_solve(int m, int n, const int64_t* col_ptr, const int* row_idx, const double* vals,
              const double* c, const double* lc, const double* uc, const double* lr,
              const double* ur, int maximize, double offset, const char* algorithm,
              int mode, int enable_presolve, int enable_scaling, double tolerance,
              double gap_tolerance, int64_t iteration_limit, int64_t check_frequency,
              int use_gpu, double* x_out, double* row_dual_out, double* col_dual_out,
              double* row_value_out, double* scal_out) {
    indus::Model model = make_model(m, n, col_ptr, row_idx, vals, c, lc, uc, lr, ur, maximize, offset);
    indus::Options opts;
    opts.algorithm = algorithm;  // set the member directly: the set(const char*) call binds to bool
    opts.enable_presolve = enable_presolve != 0;
    opts.enable_scaling = enable_scaling != 0;
    opts.use_gpu = use_gpu != 0;
    opts.iteration_limit = iteration_limit;
    opts.set("tolerance", tolerance);
    opts.set("gap_tolerance", gap_tolerance);
    opts.set("check_frequency", static_cast<int64_t>(check_frequency));
    indus::Solution sol = (mode == 0) ? indus::pdhg::solve_pdhg_cpu(model, opts)
                                      : indus::solve(model, opts);
    const auto& d = indus::pdhg::get_last_cpu_pdhg_diagnostics();
    if (x_out && sol.col_value.size() == static_cast<size_t>(n))
        std::memcpy(x_out, sol.col_value.data(), sizeof(double) * n);
    if (row_dual_out && sol.row_dual.size() == static_cast<size_t>(m))
        std::memcpy(row_dual_out, sol.row_dual.data(), sizeof(double) * m);
    if (col_dual_out && sol.col_dual.size() == static_cast<size_t>(n))
        std::memcpy(col_dual_out, sol.col_dual.data(), sizeof(double) * n);
    if (row_value_out && sol.row_value.size() == static_cast<size_t>(m))
        std::memcpy(row_value_out, sol.row_value.data(), sizeof(double) * m);
    double s[24] = {0};
    s[0] = static_cast<double>(sol.status);
    s[1] = sol.objective_value;
    s[2] = static_cast<double>(sol.iterations);
    s[3] = d.tau;
    s[4] = d.sigma;
    s[5] = d.spectral_norm_estimate;
    s[6] = d.restart_count;
    s[7] = d.primal_residual;
    s[8] = d.dual_residual;
    s[9] = d.duality_gap;
    s[10] = d.converged ? 1.0 : 0.0;
    s[11] = sol.quality.max_primal_violation;
    s[12] = sol.quality.max_dual_violation;
    s[13] = sol.quality.is_primal_feasible ? 1.0 : 0.0;
    s[14] = sol.quality.is_dual_feasible ? 1.0 : 0.0;
    s[15] = sol.quality.max_complementarity_violation;
    s[16] = sol.best_dual_bound;
    s[17] = d.setup_time_sec;
    s[18] = d.iteration_time_sec;
    s[19] = d.total_time_sec;
    s[20] = sol.presolve_total_reductions;
    std::memcpy(scal_out, s, sizeof(s));
    return 0;
}

double ref_spectral_norm(int m, int n, const int64_t* col_ptr, const int* row_idx,
                         const double* vals, int max_iter, double tol) {
    return indus::pdhg::estimate_spectral_norm(make_csc(m, n, col_ptr, row_idx, vals), max_iter, tol);
}

// In-place Ruiz on the given CSC values and vectors; writes row/col scales.
int ref_ruiz(int m, int n, const int64_t* col_ptr, const int* row_idx, double* vals,
             double* c, double* lr, double* ur, double* lc, double* uc, int max_iters,
             double tol, double* row_scale, double* col_scale) {
    auto A = make_csc(m, n, col_ptr, row_idx, vals);
    std::vector<double> vc(c, c + n), vlr(lr, lr + m), vur(ur, ur + m), vlc(lc, lc + n), vuc(uc, uc + n);
    auto sf = indus::la::RuizScaler::compute_and_scale(A, vc, vlr, vur, vlc, vuc, max_iters, tol);
    std::memcpy(vals, A.values.data(), sizeof(double) * A.values.size());
    std::memcpy(c, vc.data(), sizeof(double) * n);
    std::memcpy(lr, vlr.data(), sizeof(double) * m);
    std::memcpy(ur, vur.data(), sizeof(double) * m);
    std::memcpy(lc, vlc.data(), sizeof(double) * n);
    std::memcpy(uc, vuc.data(), sizeof(double) * n);
    if (sf.is_scaled) {
        std::memcpy(row_scale, sf.row_scale.data(), sizeof(double) * m);
        std::memcpy(col_scale, sf.col_scale.data(), sizeof(double) * n);
    }
    return sf.is_scaled ? 1 : 0;
}

// y = A x (CSC scatter with the reference's skip rule), v = A^T y_in.
int ref_spmv(int m, int n, const int64_t* col_ptr, const int* row_idx, const double* vals,
             const double* x, double* y, const double* y_in, double* v) {
    auto A = make_csc(m, n, col_ptr, row_idx, vals);
    std::vector<double> yy(static_cast<size_t>(m)), vv(static_cast<size_t>(n));
    A.multiply(std::span<const double>(x, static_cast<size_t>(n)), yy);
    A.multiply_transpose(std::span<const double>(y_in, static_cast<size_t>(m)), vv);
    std::memcpy(y, yy.data(), sizeof(double) * m);
    std::memcpy(v, vv.data(), sizeof(double) * n);
    return 0;
}

// Two-call MPS reader: first call with null arrays fills dims[0..3] = m, n, nnz, maximize.
// Model read by the reference's own MPS (lp = 0) or LP (lp = 1) reader; two calls: dims first
// (m, n, nnz, maximize, number of integer columns, Q nonzeros), then the arrays.
static int ref_read_any(int lp, const char* path, int* dims, double* offset, int64_t* col_ptr, int* row_idx,
                        double* vals, double* c, double* lc, double* uc, double* lr, double* ur) {
    indus::Model model = lp ? indus::io::read_lp(path) : indus::io::read_mps(path);
    dims[0] = model.num_rows;
    dims[1] = model.num_cols;
    dims[2] = static_cast<int>(model.A.nnz());
    dims[3] = model.sense == indus::ObjSense::kMaximize ? 1 : 0;
    if (dims[4] >= 0) {   // extended query: integer columns and Q nonzeros
        int ints = 0;
        for (auto t : model.col_type) ints += t == indus::VarType::kInteger;
        dims[4] = ints;
        dims[5] = static_cast<int>(model.Q.nnz());
    }
    *offset = model.objective_offset;
    if (!col_ptr) return 0;
    const int m = model.num_rows, n = model.num_cols;
    std::memcpy(col_ptr, model.A.col_ptr.data(), sizeof(int64_t) * (n + 1));
    std::memcpy(row_idx, model.A.row_idx.data(), sizeof(int) * model.A.nnz());
    std::memcpy(vals, model.A.values.data(), sizeof(double) * model.A.nnz());
    std::memcpy(c, model.c.data(), sizeof(double) * n);
    std::memcpy(lc, model.col_lower.data(), sizeof(double) * n);
    std::memcpy(uc, model.col_upper.data(), sizeof(double) * n);
    std::memcpy(lr, model.row_lower.data(), sizeof(double) * m);
    std::memcpy(ur, model.row_upper.data(), sizeof(double) * m);
    return 0;
}

extern "C" int ref_read_mps(const char* path, int* dims, double* offset, int64_t* col_ptr, int* row_idx,
                            double* vals, double* c, double* lc, double* uc, double* lr, double* ur) {
    return ref_read_any(0, path, dims, offset, col_ptr, row_idx, vals, c, lc, uc, lr, ur);
}

extern "C" int ref_read_lp(const char* path, int* dims, double* offset, int64_t* col_ptr, int* row_idx,
                           double* vals, double* c, double* lc, double* uc, double* lr, double* ur) {
    return ref_read_any(1, path, dims, offset, col_ptr, row_idx, vals, c, lc, uc, lr, ur);
}

// verify_solution on a caller-given point; out: 0 passed, 1 max_primal_violation,
// 2 max_dual_violation, 3 max_complementarity_violation, 4 objective_discrepancy,
// 5 recomputed_objective, 6 primal_feasible, 7 dual_feasible, 8 bounds_feasible, 9 objective_matches
int ref_verify(int m, int n, const int64_t* col_ptr, const int* row_idx, const double* vals,
               const double* c, const double* lc, const double* uc, const double* lr,
               const double* ur, int maximize, double offset, const double* x,
               const double* row_dual, double objective, int status, double tol, double* out) {
    indus::Model model = make_model(m, n, col_ptr, row_idx, vals, c, lc, uc, lr, ur, maximize, offset);
    indus::Solution sol;
    sol.status = static_cast<indus::SolveStatus>(status);
    sol.objective_value = objective;
    sol.col_value.assign(x, x + n);
    if (row_dual) sol.row_dual.assign(row_dual, row_dual + m);
    auto r = indus::verifier::verify_solution(model, sol, tol);
    out[0] = r.passed; out[1] = r.max_primal_violation; out[2] = r.max_dual_violation;
    out[3] = r.max_complementarity_violation; out[4] = r.objective_discrepancy;
    out[5] = r.recomputed_objective; out[6] = r.primal_feasible; out[7] = r.dual_feasible;
    out[8] = r.bounds_feasible; out[9] = r.objective_matches;
    return 0;
}

// Solution::recompute_quality on a caller-given point (status OPTIMAL).
// out: 0 max_primal_violation, 1 max_dual_violation, 2 max_complementarity_violation,
// 3 is_primal_feasible, 4 is_dual_feasible
int ref_quality(int m, int n, const int64_t* col_ptr, const int* row_idx, const double* vals,
                const double* c, const double* lc, const double* uc, const double* lr,
                const double* ur, int maximize, double offset, const double* x,
                const double* row_dual, double* col_dual_out, double* out) {
    indus::Model model = make_model(m, n, col_ptr, row_idx, vals, c, lc, uc, lr, ur, maximize, offset);
    indus::Solution sol;
    sol.status = indus::SolveStatus::kOptimal;
    sol.col_value.assign(x, x + n);
    sol.row_dual.assign(row_dual, row_dual + m);
    sol.recompute_quality(model);
    out[0] = sol.quality.max_primal_violation; out[1] = sol.quality.max_dual_violation;
    out[2] = sol.quality.max_complementarity_violation;
    out[3] = sol.quality.is_primal_feasible; out[4] = sol.quality.is_dual_feasible;
    if (col_dual_out) std::memcpy(col_dual_out, sol.col_dual.data(), sizeof(double) * n);
    return 0;
}

// Presolve of a model file read by the reference's reader, then the postsolve of the
// synthetic reduced solution, both written as the canonical dump of presolve_dump.hpp.
int ref_presolve_dump(const char* path, int lp, const char* out_path) {
    indus::Model model = lp ? indus::io::read_lp(path) : indus::io::read_mps(path);
    auto res = indus::presolve::PresolveEngine::apply(model);
    FILE* f = std::fopen(out_path, "w");
    if (!f) return 1;
    presolve_dump::dump_result(f, res);
    if (!res.is_infeasible && !res.is_unbounded) {
        auto red = presolve_dump::synthetic_reduced<indus::Solution>(res.presolved_model, indus::SolveStatus::kOptimal);
        presolve_dump::dump_solution(f, indus::presolve::PresolveEngine::postsolve(red, res.stack, model));
    }
    std::fclose(f);
    return 0;
}

// The audits of audit_dump.hpp (verify_solution, recompute_quality, verify_solution) on the
// point in point_path for the model in the MPS file model_path, written to out_path.
int ref_audit_dump(const char* model_path, const char* point_path, double tol, const char* out_path) {
    const indus::Model model = indus::io::read_mps(model_path);
    audit_dump::Point p;
    if (!audit_dump::read_point(point_path, p)) return 1;
    FILE* f = std::fopen(out_path, "w");
    if (!f) return 1;
    audit_dump::run<indus::Solution, indus::Model, indus::SolveStatus>(
        f, model, p, tol, [](const indus::Model& m, const indus::Solution& s, double t) {
            return indus::verifier::verify_solution(m, s, t);
        });
    std::fclose(f);
    return 0;
}

// The reference's .sol and JSON writers on the Solution audit_dump::solution_of builds from
// the point in point_path (src/io/writer.cpp:85-264).
int ref_write_files(const char* model_path, const char* point_path, const char* sol_path, const char* json_path) {
    const indus::Model model = indus::io::read_mps(model_path);
    audit_dump::Point p;
    if (!audit_dump::read_point(point_path, p)) return 1;
    const auto s = audit_dump::solution_of<indus::Solution, indus::Model, indus::SolveStatus>(model, p);
    indus::io::write_solution(s, model, sol_path);
    indus::io::write_json(s, model, json_path);
    return 0;
}

// verify_files (src/engine/verifier.cpp:660-668: the model file read by the reference's
// reader, the .sol parsed by its parser, then verify_solution), written as the verification
// dump of audit_dump.hpp; returns 1 when the verification passed, 0 when not, -1 on I/O error.
int ref_verify_files(const char* model_path, const char* sol_path, double tol, const char* out_path) {
    const auto r = indus::verifier::verify_files(model_path, sol_path, tol);
    FILE* f = std::fopen(out_path, "w");
    if (!f) return -1;
    audit_dump::verification(f, "files", r);
    std::fclose(f);
    return r.passed ? 1 : 0;
}

// The options probe script of options_probe.hpp run on the reference's Options; returns the
// dump's length (the text is cut at cap - 1 bytes).
int ref_options_probe(char* out, int cap) {
    const std::string s = options_probe::run<indus::Options>();
    if (cap > 0) {
        const size_t k = std::min(s.size(), static_cast<size_t>(cap - 1));
        std::memcpy(out, s.data(), k);
        out[k] = 0;
    }
    return static_cast<int>(s.size());
}

}  // extern "C"
