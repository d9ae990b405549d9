// Test driver of the C++ solver API (include/indus/*.hpp). It runs the reference's PDHG test
// cases (tests/test_pdhg.cpp:54-436 of the reference) through indus::solve, plus checks of
// the MPS reader, the writers, the verifier, the presolve handoff and the batch entry point.
//
//   test_indus_api cpu <data_dir>   CPU PDHG (algorithm "pdhg_cpu"); no device needed
//   test_indus_api gpu <data_dir>   MI355X engine (algorithm "pdhg_cuda")
//
//   test_indus_api dump <file.mps> <iters> <tol> <cpu|gpu-reference|gpu>
//   test_indus_api presolve_scale <m> <n>
//   test_indus_api dump_model <file> <mps|lp>
//
// data_dir holds afiro.mps and crude_blend.mps. Exit status 0 iff every check passed.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <functional>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "indus/gpu.hpp"
#include "indus/io.hpp"
#include "indus/model.hpp"
#include "indus/pdhg.hpp"
#include "indus/presolve.hpp"
#include "indus/verifier.hpp"
#include "audit_dump.hpp"
#include "options_probe.hpp"
#include "presolve_dump.hpp"
#include "vajra_lpgen.h"

namespace {

int g_fail = 0;
bool g_gpu = false;
std::string g_dir;

#define CHECK(cond)                                                                          \
    do {                                                                                     \
        if (!(cond)) {                                                                       \
            std::cout << "  CHECK FAILED " << __FILE__ << ":" << __LINE__ << ": " #cond "\n"; \
            ++g_fail;                                                                        \
        }                                                                                    \
    } while (0)
#define CHECK_NEAR(a, b, tol)                                                                                 \
    do {                                                                                                      \
        const double a_ = (a), b_ = (b);                                                                      \
        if (!(std::fabs(a_ - b_) <= (tol))) {                                                                 \
            std::cout << "  CHECK_NEAR FAILED " << __FILE__ << ":" << __LINE__ << ": " #a " = " << a_ << " vs " \
                      << b_ << " (tol " << (tol) << ")\n";                                                    \
            ++g_fail;                                                                                         \
        }                                                                                                     \
    } while (0)

// The reference's test cases ask for tolerance 1e-5, but its dispatcher's status guard
// (model.cpp:626-664) audits the answer at 1e-7 and downgrades a 1e-5 PDHG point (checked in
// test_status_guard). The KATs therefore ask for 1e-7 on the CPU and a relative KKT error of
// 1e-8 on the device.
indus::Options pdhg_options(double tol = 0.0, int64_t iters = 20000) {
    if (tol == 0.0) tol = g_gpu ? 1e-8 : 1e-7;
    indus::Options o;
    if (g_gpu) {
        o.set("algorithm", "pdhg_cuda");
        o.use_gpu = true;
    } else {
        o.set("algorithm", "pdhg_cpu");
    }
    o.enable_presolve = false;
    o.set("tolerance", tol);
    o.iteration_limit = iters;
    return o;
}

indus::Model make(int m, int n, std::vector<double> c, std::vector<double> lc, std::vector<double> uc,
                  std::vector<double> lr, std::vector<double> ur, const std::vector<indus::la::Triplet>& t) {
    indus::Model M;
    M.num_rows = m;
    M.num_cols = n;
    M.c = std::move(c);
    M.col_lower = std::move(lc);
    M.col_upper = std::move(uc);
    M.row_lower = std::move(lr);
    M.row_upper = std::move(ur);
    M.A.set_from_triplets(m, n, t);
    return M;
}

indus::Model two_var() {
    return make(3, 2, {-1, -2}, {0, 0}, {1e20, 1e20}, {-1e20, -1e20, -1e20}, {4, 3, 3},
                {{0, 0, 1.0}, {0, 1, 1.0}, {1, 0, 1.0}, {2, 1, 1.0}});
}

void verified(const indus::Model& M, const indus::Solution& s, double tol = 1e-4) {
    const auto v = indus::verifier::verify_solution(M, s, tol);
    if (!v.passed) std::cout << "  " << v.summary << "\n";
    CHECK(v.passed);
}

void test_two_var() {
    const indus::Model M = two_var();
    const indus::Solution s = indus::solve(M, pdhg_options());
    CHECK(s.status == indus::SolveStatus::kOptimal);
    CHECK_NEAR(s.objective_value, -7.0, 1e-3);
    CHECK_NEAR(s.col_value[0], 1.0, 1e-3);
    CHECK_NEAR(s.col_value[1], 3.0, 1e-3);
    CHECK(s.algorithm_used == (g_gpu ? "pdhg_cuda" : "pdhg_cpu"));
    verified(M, s);
}

void test_equality() {
    const indus::Model M = make(2, 3, {1, 2, 3}, {0, 0, 0}, {1e20, 1e20, 1e20}, {6, 5}, {6, 5},
                                {{0, 0, 1.0}, {0, 1, 1.0}, {0, 2, 1.0}, {1, 0, 2.0}, {1, 1, 1.0}});
    const indus::Solution s = indus::solve(M, pdhg_options());
    CHECK(s.status == indus::SolveStatus::kOptimal);
    CHECK_NEAR(s.objective_value, 13.0, 1e-3);
    CHECK_NEAR(s.col_value[0] + s.col_value[1] + s.col_value[2], 6.0, 1e-3);
    CHECK_NEAR(2.0 * s.col_value[0] + s.col_value[1], 5.0, 1e-3);
    verified(M, s);
}

void test_mixed_senses() {
    const indus::Model M = make(3, 2, {2, 3}, {0, 0}, {10, 10}, {2, -1e20, 1}, {1e20, 4, 6},
                                {{0, 0, 1.0}, {0, 1, 1.0}, {1, 0, 1.0}, {1, 1, -1.0}, {2, 0, 2.0}, {2, 1, 1.0}});
    const indus::Solution s = indus::solve(M, pdhg_options());
    CHECK(s.status == indus::SolveStatus::kOptimal);
    CHECK_NEAR(s.objective_value, 4.0, 1e-3);
    verified(M, s);
}

void test_free_variable() {
    const indus::Model M = make(1, 2, {2, 1}, {-10, -1e20}, {10, 1e20}, {5}, {5}, {{0, 0, 1.0}, {0, 1, 1.0}});
    const indus::Solution s = indus::solve(M, pdhg_options());
    CHECK(s.status == indus::SolveStatus::kOptimal);
    CHECK_NEAR(s.objective_value, -5.0, 1e-3);
    CHECK_NEAR(s.col_value[0], -10.0, 1e-3);
    CHECK_NEAR(s.col_value[1], 15.0, 1e-3);
    verified(M, s);
}

void test_min_max_symmetry() {
    indus::Model mn = make(1, 2, {-2, -3}, {0, 0}, {1e20, 1e20}, {-1e20}, {5}, {{0, 0, 1.0}, {0, 1, 1.0}});
    indus::Model mx = mn;
    mx.sense = indus::ObjSense::kMaximize;
    mx.c = {2, 3};
    const indus::Solution a = indus::solve(mn, pdhg_options()), b = indus::solve(mx, pdhg_options());
    CHECK(a.status == indus::SolveStatus::kOptimal);
    CHECK(b.status == indus::SolveStatus::kOptimal);
    CHECK_NEAR(a.objective_value, -15.0, 1e-3);
    CHECK_NEAR(b.objective_value, 15.0, 1e-3);
    CHECK_NEAR(a.col_value[1], b.col_value[1], 1e-3);
    // one dual convention for both senses: the <= row prices at -3 in the canonical minimisation
    CHECK_NEAR(a.row_dual[0], b.row_dual[0], 1e-3);
    CHECK_NEAR(a.row_dual[0], -3.0, 1e-3);
    verified(mx, b);
}

void test_presolve_handoff() {
    const indus::Model M = make(2, 3, {-1, -1, 2}, {0, 0, 5}, {1e20, 1e20, 5}, {-1e20, -1e20}, {4, 10},
                                {{0, 0, 1.0}, {0, 1, 1.0}, {1, 1, 1.0}});
    const auto pr = indus::presolve::PresolveEngine::apply(M);
    // x3 (fixed at 5) has no entries: the empty-column step, which runs before the
    // fixed-column step, takes it; the row x2 <= 10 becomes a bound (presolve.cpp:109-245)
    CHECK(pr.num_empty_cols == 1 && pr.num_fixed_cols == 0);
    CHECK(pr.num_singleton_rows == 1);
    CHECK(pr.total_reductions() == 2);
    CHECK(pr.presolved_model.num_cols == 2 && pr.presolved_model.num_rows == 1);
    indus::Options o = pdhg_options();
    o.enable_presolve = true;
    const indus::Solution s = indus::solve(M, o);
    CHECK(s.status == indus::SolveStatus::kOptimal);
    CHECK_NEAR(s.objective_value, 6.0, 1e-3);
    CHECK_NEAR(s.col_value[2], 5.0, 1e-5);
    CHECK(s.presolve_total_reductions == pr.total_reductions());
    CHECK(static_cast<int>(s.col_value.size()) == 3 && static_cast<int>(s.row_dual.size()) == 2);
    verified(M, s);
}

void test_status_guard() {
    // a 1e-5 CPU PDHG point of the two-variable LP violates a row by ~1e-6 > 1e-7: the guard
    // turns OPTIMAL into NUMERICAL_ERROR, as the reference's dispatcher does
    indus::Options o = pdhg_options(1e-5);
    o.set("algorithm", "pdhg_cpu");
    o.use_gpu = false;
    const indus::Solution s = indus::solve(two_var(), o);
    CHECK(s.status == indus::SolveStatus::kNumericalError);
    CHECK(s.status_message.find("Status downgraded to NUMERICAL_ERROR: primal feasibility violation") == 0);
    CHECK_NEAR(s.objective_value, -7.0, 1e-3);
}

void test_infeasible() {
    const indus::Model M = make(2, 2, {1, 1}, {0, 0}, {10, 10}, {-1e20, 3}, {1, 1e20},
                                {{0, 0, 1.0}, {0, 1, 1.0}, {1, 0, 1.0}, {1, 1, 1.0}});
    indus::Options o = pdhg_options(1e-6, 500);
    const indus::Solution s = indus::solve(M, o);
    CHECK(s.status != indus::SolveStatus::kOptimal);
}

void test_iteration_limit() {
    const indus::Model M = indus::io::read_mps(g_dir + "/afiro.mps");
    const indus::Solution s = indus::solve(M, pdhg_options(1e-6, 10));
    CHECK(s.status == indus::SolveStatus::kIterationLimit || s.status == indus::SolveStatus::kFeasible);
    CHECK(s.status != indus::SolveStatus::kOptimal);
}

void test_netlib() {
    // published optima: afiro -464.753142857, crude_blend 214.145945946
    const char* files[2] = {"afiro.mps", "crude_blend.mps"};
    const double opt[2] = {-464.75314285714285, 214.14594594594598};
    for (int q = 0; q < 2; ++q) {
        const indus::Model M = indus::io::read_mps(g_dir + "/" + files[q]);
        // device: the status guard audits every row at 1e-7 absolute or relative to the row's
        // own bound, which a relative KKT error of 1e-8 (norms over all rows) does not imply
        // on afiro; 1e-10 does
        indus::Options o = pdhg_options(g_gpu ? 1e-10 : 1e-4, g_gpu ? 200000 : 50000);
        o.enable_presolve = true;
        const indus::Solution s = indus::solve(M, o);
        std::cout << "  status " << indus::to_string(s.status) << ": " << s.status_message << "\n";
        // CPU: the reference's 1e-4 run is downgraded by the 1e-7 status guard
        CHECK(s.status == indus::SolveStatus::kOptimal ||
              (!g_gpu && s.status_message.find("Status downgraded") == 0));
        const double rel = std::fabs(s.objective_value - opt[q]) / (1.0 + std::fabs(opt[q]));
        std::cout << "  " << M.name << ": obj " << s.objective_value << " rel diff " << rel << " iterations "
                  << s.iterations << "\n";
        CHECK(rel < (g_gpu ? 1e-6 : 1e-3));
        if (g_gpu) {
            const auto k = indus::verifier::relative_kkt(M, s);
            std::cout << "    rel-KKT " << k.rel_kkt << "\n";
            CHECK(k.rel_kkt <= 1e-10 * 1.0000001);
        }
    }
}

void test_device_and_fallback() {
    const auto info = indus::gpu::probe_device();
    std::cout << "  device: " << info.device_name << " available " << info.available << "\n";
    CHECK(info.available == g_gpu);
    CHECK(indus::gpu::is_cuda_available() == g_gpu);
    const indus::Model M = make(1, 2, {1, 1}, {0, 0}, {10, 10}, {3}, {10}, {{0, 0, 1.0}, {0, 1, 1.0}});
    indus::Options o;
    o.set("algorithm", "pdhg_cuda");
    o.use_gpu = true;
    o.enable_presolve = false;
    o.iteration_limit = 5000;
    // the device solves to this relative KKT error; at the default 1e-6 whether the returned
    // point also passes the reference's 1e-7 status guard depends on where the iteration
    // stops (the guard itself is covered by status_guard)
    o.set("tolerance", 1e-9);
    if (g_gpu) {
        const indus::Solution s = indus::solve(M, o);
        CHECK(s.status == indus::SolveStatus::kOptimal);
        CHECK_NEAR(s.objective_value, 3.0, 1e-3);
        const auto t = indus::gpu::get_last_gpu_timing();
        std::cout << "  kernel mode: " << t.kernel_mode << "\n";
        CHECK(t.kernel_mode.find("restarted_halpern") == 0);
        CHECK(t.iterations == s.iterations);
        CHECK(t.rel_kkt <= 1e-9);
    } else {
        // no device: a device request fails loudly ...
        bool threw = false;
        try {
            (void)indus::solve(M, o);
        } catch (const std::runtime_error&) {
            threw = true;
        }
        CHECK(threw);
        // ... unless the caller asks for the reference's CPU fallback (gpu_context.cpp:65-71)
        o.set("allow_cpu_fallback", true);
        o.set("tolerance", 1e-7);
        const indus::Solution s = indus::solve(M, o);
        CHECK(s.status == indus::SolveStatus::kOptimal);
        CHECK_NEAR(s.objective_value, 3.0, 1e-3);
        CHECK(s.algorithm_used == "pdhg_cpu");
        CHECK(indus::gpu::get_last_gpu_timing().kernel_mode == "CPU reference engine");
    }
}

void test_mps_reader() {
    const indus::Model M = indus::io::read_mps(g_dir + "/afiro.mps");
    CHECK(M.name == "AFIRO");
    CHECK(M.num_rows == 27 && M.num_cols == 32);
    CHECK(M.A.nnz() == 83);
    CHECK(M.sense == indus::ObjSense::kMinimize);
    CHECK(M.classify() == indus::ProblemClass::kLp);
    CHECK(M.row_names[0] == "R09" && M.col_names[0] == "X01");
    bool threw = false;
    try {
        (void)indus::io::read_mps(g_dir + "/no_such_file.mps");
    } catch (const std::runtime_error&) {
        threw = true;
    }
    CHECK(threw);
    // every quirk of the format reader on one small file
    const std::string p = "/tmp/indus_quirks_test.mps";
    {
        std::ofstream f(p);
        f << "NAME QUIRKS\n* comment\nOBJSENSE\n    MAX\nROWS\n N obj\n E e1\n L l1\n G g1\n E e2\n"
             "COLUMNS\n    x obj 1.5D0 e1 1\n    x l1 2 g1 3\n    MARKER 'MARKER' 'INTORG'\n    y obj 1 e2 1\n"
             "    MARKER 'MARKER' 'INTEND'\n    z e1 1e-12 l1 1\nRHS\n    rhs obj 7 e1 4\n    rhs l1 5\n    g1 2\n"
             "    rhs e2 3\nRANGES\n    rng e1 2 l1 3\n    rng g1 -1\n    rng e2 -2\nBOUNDS\n UP bnd x -4\n FR bnd y\n"
             " FX bnd z 2.5\nENDATA\n";
    }
    const indus::Model Q = indus::io::read_mps(p);
    std::remove(p.c_str());
    CHECK(Q.name == "QUIRKS");
    CHECK(Q.sense == indus::ObjSense::kMaximize);
    CHECK(Q.num_rows == 4 && Q.num_cols == 3);
    CHECK_NEAR(Q.c[0], 1.5, 0.0);
    CHECK(Q.objective_offset == 0.0);  // the objective row's RHS is ignored
    CHECK(Q.A.nnz() == 5);             // 1e-12 dropped
    CHECK(Q.col_type[1] == indus::VarType::kInteger && Q.col_type[0] == indus::VarType::kContinuous);
    CHECK(Q.row_lower[0] == 4 && Q.row_upper[0] == 6);   // E, r > 0
    CHECK(Q.row_lower[1] == 2 && Q.row_upper[1] == 5);   // L, b - |r|
    CHECK(Q.row_lower[2] == 2 && Q.row_upper[2] == 3);   // G, b + |r| (two-token RHS line)
    CHECK(Q.row_lower[3] == 1 && Q.row_upper[3] == 3);   // E, r < 0
    CHECK(Q.col_lower[0] == -1e20 && Q.col_upper[0] == -4);
    CHECK(Q.col_lower[1] == -1e20 && Q.col_upper[1] == 1e20);
    CHECK(Q.col_lower[2] == 2.5 && Q.col_upper[2] == 2.5);
}

// writes M in LP format (every row as "lo <= expr <= hi" or one-sided) for a round trip
void write_lp_file(const indus::Model& M, const std::string& path) {
    std::ofstream f(path);
    f << std::setprecision(17);
    f << (M.sense == indus::ObjSense::kMaximize ? "Maximize\n" : "Minimize\n") << " obj:";
    for (int j = 0; j < M.num_cols; ++j) f << (M.c[j] < 0 ? " - " : " + ") << std::fabs(M.c[j]) << " " << M.col_names[j];
    f << "\nSubject To\n";
    const indus::la::SparseMatrixCSR R = M.A.to_csr();
    for (int i = 0; i < M.num_rows; ++i) {
        f << " " << M.row_names[i] << ": ";
        const bool lo = M.row_lower[i] > -1e20, hi = M.row_upper[i] < 1e20;
        if (lo && hi && M.row_lower[i] != M.row_upper[i]) f << M.row_lower[i] << " <= ";
        for (int64_t k = R.row_ptr[i]; k < R.row_ptr[i + 1]; ++k)
            f << (R.values[k] < 0 ? " - " : " + ") << std::fabs(R.values[k]) << " " << M.col_names[R.col_idx[k]];
        if (lo && hi && M.row_lower[i] == M.row_upper[i]) f << " = " << M.row_upper[i] << "\n";
        else if (hi) f << " <= " << M.row_upper[i] << "\n";
        else f << " >= " << M.row_lower[i] << "\n";
    }
    f << "Bounds\n";
    for (int j = 0; j < M.num_cols; ++j) {
        const double l = M.col_lower[j], u = M.col_upper[j];
        if (l <= -1e20 && u >= 1e20) f << " " << M.col_names[j] << " free\n";
        else
            f << " " << (l <= -1e20 ? -1e30 : l) << " <= " << M.col_names[j] << " <= " << (u >= 1e20 ? 1e30 : u) << "\n";
    }
    f << "End\n";
}

void test_lp_reader() {
    // afiro through an LP-format round trip reads into the same LP
    const indus::Model A = indus::io::read_mps(g_dir + "/afiro.mps");
    const std::string path = "/tmp/indus_afiro_roundtrip.lp";
    write_lp_file(A, path);
    const indus::Model B = indus::io::read_lp(path);
    std::remove(path.c_str());
    CHECK(B.num_rows == A.num_rows && B.num_cols == A.num_cols && B.A.nnz() == A.A.nnz());
    CHECK(B.name == "obj");
    bool same = B.A.col_ptr == A.A.col_ptr && B.A.row_idx == A.A.row_idx && B.A.values == A.A.values && B.c == A.c &&
                B.row_lower == A.row_lower && B.row_upper == A.row_upper;
    CHECK(same);
    for (int j = 0; j < A.num_cols; ++j) {
        CHECK(B.col_lower[j] == (A.col_lower[j] <= -1e20 ? -1e30 : A.col_lower[j]));
        CHECK(B.col_upper[j] == (A.col_upper[j] >= 1e20 ? 1e30 : A.col_upper[j]));
    }
    // the grammar's corners
    const std::string q = "/tmp/indus_quirks_test.lp";
    {
        std::ofstream f(q);
        f << "\\ comment line\nMAXIMIZE\n profit: 3 x + 2y - z [ 2 x ^2 + 4 x * y ] / 2\nS.T.\n"
             " c1: x + y <= 4\n 2 x - y >= -2\n r3: -1 <= x - z <= 5.5\n c4: y = 1.5e0\n c5: x + z > 1\n"
             "Bounds\n y <= 10\n -inf <= z <= 3\n x >= -2\n w free\n 1 <= v <= INF\n"
             "General\n v\nBinary\n b\nEnd\n ignored: q + r <= 1\n";
    }
    const indus::Model L = indus::io::read_lp(q);
    std::remove(q.c_str());
    CHECK(L.sense == indus::ObjSense::kMaximize);
    CHECK(L.name == "profit");
    CHECK(L.num_rows == 5);
    CHECK(L.row_names[1] == "R1");
    CHECK(L.row_lower[0] == -1e20 && L.row_upper[0] == 4);
    CHECK(L.row_lower[1] == -2 && L.row_upper[1] == 1e20);
    CHECK(L.row_lower[2] == -1 && L.row_upper[2] == 5.5);
    CHECK(L.row_lower[3] == 1.5 && L.row_upper[3] == 1.5);
    CHECK(L.row_lower[4] == 1 && L.row_upper[4] == 1e20);
    // columns in order of appearance: x, y, z, w, v, b
    CHECK(L.num_cols == 6 && L.col_names[0] == "x" && L.col_names[1] == "y" && L.col_names[2] == "z");
    CHECK(L.c[0] == 3 && L.c[1] == 2 && L.c[2] == -1);
    CHECK(L.Q.nnz() == 2);
    CHECK(L.col_lower[1] == 0 && L.col_upper[1] == 10);
    CHECK(L.col_lower[2] == -1e20 && L.col_upper[2] == 3);
    CHECK(L.col_lower[0] == -2 && L.col_upper[0] == 1e20);
    CHECK(L.col_lower[3] == -1e20 && L.col_upper[3] == 1e20);
    CHECK(L.col_lower[4] == 1 && L.col_upper[4] == 1e20 && L.col_type[4] == indus::VarType::kInteger);
    CHECK(L.col_lower[5] == 0 && L.col_upper[5] == 1 && L.col_type[5] == indus::VarType::kInteger);
}

void test_writers() {
    const indus::Model M = two_var();
    indus::Options o = pdhg_options();
    const indus::Solution s = indus::solve(M, o);
    const std::string sol = "/tmp/indus_writer_test.sol", js = "/tmp/indus_writer_test.json";
    indus::io::write_solution(s, M, sol);
    indus::io::write_json(s, M, js);
    std::ifstream f(sol);
    std::string line;
    int cols = 0, rows = 0, section = 0;
    bool status_ok = false;
    while (std::getline(f, line)) {
        if (line == "# Status: OPTIMAL") status_ok = true;
        if (line == "# Columns (Variables)") section = 1;
        if (line == "# Rows (Constraints)") section = 2;
        if (line.empty() || line[0] == '#') continue;
        std::istringstream ls(line);
        std::string name, st;
        double v, d;
        ls >> name >> v >> d >> st;
        if (section == 1) {
            CHECK(v == s.col_value[cols]);  // 17 digits round-trip exactly
            ++cols;
        } else if (section == 2) {
            CHECK(d == s.row_dual[rows]);
            ++rows;
        }
    }
    CHECK(status_ok);
    CHECK(cols == 2 && rows == 3);
    std::ifstream g(js);
    std::stringstream all;
    all << g.rdbuf();
    const std::string j = all.str();
    CHECK(j.find("\"verification_status\": \"PASSED\"") != std::string::npos);
    CHECK(j.find("{\"name\": \"r2\", \"activity\": ") != std::string::npos);  // unnamed rows print as r<i>
    CHECK(j.find("{\"name\": \"c1\", \"value\": ") != std::string::npos);
    CHECK(j.find("\"constraints\": [") != std::string::npos);
    std::remove(sol.c_str());
    std::remove(js.c_str());
}

void test_verifier_rejects() {
    const indus::Model M = two_var();
    indus::Solution s;
    s.status = indus::SolveStatus::kOptimal;
    s.col_value = {2.0, 3.0};  // x1 + x2 = 5 > 4
    s.objective_value = -8.0;
    const auto v = indus::verifier::verify_solution(M, s, 1e-6);
    CHECK(!v.passed && !v.primal_feasible);
    CHECK_NEAR(v.max_row_violation, 1.0, 1e-12);
    s.col_value = {1.0, 3.0};
    s.objective_value = -7.5;  // wrong objective
    CHECK(!indus::verifier::verify_solution(M, s, 1e-6).objective_matches);
    s.objective_value = -7.0;
    s.row_dual = {-1.0, 0.0, -1.0};  // exact optimal duals of the canonical minimisation
    const auto ok = indus::verifier::verify_solution(M, s, 1e-9);
    CHECK(ok.passed);
    const auto k = indus::verifier::relative_kkt(M, s);
    CHECK_NEAR(k.rel_kkt, 0.0, 1e-15);
}

void test_cpu_gpu_reference_map_identical() {
    // the device reference map is the CPU PDHG's iteration map in the same operation order:
    // same iterates, same iteration count, same answer bit for bit
    const indus::Model M = indus::io::read_mps(g_dir + "/afiro.mps");
    indus::Options o = pdhg_options(1e-6, 3000);
    indus::Options c = o;
    c.set("algorithm", "pdhg_cpu");
    c.use_gpu = false;
    o.set("pdhg_mode", "reference");
    const indus::Solution a = indus::pdhg::solve_pdhg_cpu(M, c);
    const indus::Solution b = indus::gpu::solve_pdhg_gpu(M, o);
    CHECK(a.iterations == b.iterations);
    CHECK(a.status == b.status);
    bool same = a.col_value.size() == b.col_value.size();
    for (size_t j = 0; same && j < a.col_value.size(); ++j) same = a.col_value[j] == b.col_value[j];
    for (size_t i = 0; same && i < a.row_dual.size(); ++i) same = a.row_dual[i] == b.row_dual[i];
    CHECK(same);
    CHECK(indus::gpu::get_last_gpu_timing().kernel_mode == "reference_map_exact_order");
}

void test_batch() {
    std::vector<indus::Model> ms = {two_var(), make(1, 2, {2, 1}, {-10, -1e20}, {10, 1e20}, {5}, {5},
                                                    {{0, 0, 1.0}, {0, 1, 1.0}})};
    indus::Model mx = make(1, 2, {2, 3}, {0, 0}, {1e20, 1e20}, {-1e20}, {5}, {{0, 0, 1.0}, {0, 1, 1.0}});
    mx.sense = indus::ObjSense::kMaximize;
    ms.push_back(mx);
    ms.push_back(indus::io::read_mps(g_dir + "/afiro.mps"));
    // objective offsets enter the relative gap of the stopping test: one per sense
    indus::Model off_min = two_var();
    off_min.objective_offset = 1000.0;
    ms.push_back(off_min);
    indus::Model off_max = mx;
    off_max.objective_offset = -50.0;
    ms.push_back(off_max);
    indus::Options o = pdhg_options(1e-8, 100000);
    const auto sols = indus::gpu::solve_pdhg_gpu_batch(ms, o);
    const double opt[6] = {-7.0, -5.0, 15.0, -464.75314285714285, 993.0, -35.0};
    CHECK(sols.size() == 6);
    for (size_t q = 0; q < sols.size(); ++q) {
        CHECK(sols[q].status == indus::SolveStatus::kOptimal);
        CHECK_NEAR(sols[q].objective_value, opt[q], 1e-6 * (1 + std::fabs(opt[q])));
        CHECK(indus::verifier::relative_kkt(ms[q], sols[q]).rel_kkt <= 1e-8 * 1.0000001);
    }
}

// time_limit <= 0 answers TIME_LIMIT before any solve, with the reference's message
// (model.cpp:403-409), on the CPU and the device path alike.
void test_time_limit_zero() {
    const indus::Model M = two_var();
    for (double tl : {0.0, -1.0}) {
        indus::Options o = pdhg_options(1e-6, 1000);
        o.time_limit = tl;
        const indus::Solution s = indus::solve(M, o);
        CHECK(s.status == indus::SolveStatus::kTimeLimit);
        CHECK(s.status_message == "Time limit exceeded before solve commenced");
        CHECK(s.iterations == 0 && s.col_value.empty());
    }
}

// Models without rows, or with rows but no nonzeros, go to the CPU PDHG, as the reference's
// GPU entry delegates them (pdhg_cuda.cu:127-134): the same solution as solve_pdhg_cpu.
void test_gpu_trivial_sizes() {
    const std::vector<indus::Model> ms = {
        make(0, 3, {1, -1, 0}, {0, 0, -1}, {2, 2, 1}, {}, {}, {}),
        make(2, 3, {1, -1, 0}, {0, 0, -1}, {2, 2, 1}, {-1, -1e20}, {1, 4}, {})};
    const indus::Options o = pdhg_options(1e-6, 1000);
    for (const auto& M : ms) {
        const indus::Solution g = indus::gpu::solve_pdhg_gpu(M, o);
        const indus::Solution c = indus::pdhg::solve_pdhg_cpu(M, o);
        CHECK(g.status == c.status);
        CHECK(g.objective_value == c.objective_value);
        CHECK(g.col_value == c.col_value);
        CHECK(g.row_dual == c.row_dual);
        CHECK_NEAR(g.objective_value, -2.0, 1e-9);
    }
}

void test_mixed_precision_option() {
    const indus::Model M = indus::io::read_mps(g_dir + "/crude_blend.mps");
    indus::Options o = pdhg_options(1e-8, 200000);
    for (const char* prec : {"fp32", "bf16", "fp16"}) {
        // 0: the per-precision default; -1 / 1: the FP64 centre moves at restarts only / at every check
        for (int refine : {0, -1, 1}) {
            o.set("pdhg_precision", prec);
            o.set("pdhg_refine_every_check", refine);
            const indus::Solution s = indus::solve(M, o);
            CHECK(s.status == indus::SolveStatus::kOptimal);
            CHECK_NEAR(s.objective_value, 214.14594594594598, 1e-6 * 215.0);
            const auto t = indus::gpu::get_last_gpu_timing();
            CHECK(t.kernel_mode.find(std::string(prec) + "_refined") != std::string::npos);
            std::cout << "  " << prec << " refine_every_check " << refine << ": iterations " << s.iterations
                      << " mode " << t.kernel_mode << "\n";
        }
    }
}

}  // namespace

// dump <file.mps> <iteration_limit> <tolerance> <cpu|gpu-reference|gpu>: one PDHG solve
// without presolve; prints status, iterations and the answer in hexadecimal floating point,
// so a caller can compare it bit for bit with another implementation.
int dump(int argc, char** argv) {
    if (argc < 6) return 2;
    const indus::Model M = indus::io::read_mps(argv[2]);
    const std::string how = argv[5];
    indus::Options o;
    o.enable_presolve = false;
    o.iteration_limit = std::stoll(argv[3]);
    o.set("tolerance", std::stod(argv[4]));
    indus::Solution s;
    if (how == "cpu") {
        o.set("algorithm", "pdhg_cpu");
        indus::pdhg::CpuPdhgSolver solver(M, o);
        s = solver.solve();
        const auto& d = solver.diagnostics();
        std::printf("diag %lld %d %a %a %a\n", static_cast<long long>(d.iterations), d.restart_count, d.tau, d.sigma,
                    d.spectral_norm_estimate);
    } else {
        if (how == "gpu-reference") o.set("pdhg_mode", "reference");
        s = indus::gpu::solve_pdhg_gpu(M, o);
    }
    std::printf("status %d\niterations %lld\nobjective %a\nx", static_cast<int>(s.status),
                static_cast<long long>(s.iterations), s.objective_value);
    for (double v : s.col_value) std::printf(" %a", v);
    std::printf("\nrow_dual");
    for (double v : s.row_dual) std::printf(" %a", v);
    std::printf("\ncol_dual");
    for (double v : s.col_dual) std::printf(" %a", v);
    std::printf("\nrow_value");
    for (double v : s.row_value) std::printf(" %a", v);
    std::printf("\n");
    return 0;
}

// presolve_scale <m> <n>: the presolve -> postsolve handoff on a large random LP with
// planted reductions (every 50th column fixed at its feasible value, 1000 empty rows, 1000
// singleton rows); prints the timings and checks that postsolve restores the feasible point.
int presolve_scale(int argc, char** argv) {
    if (argc < 4) return 2;
    const int m = std::atoi(argv[2]), n = std::atoi(argv[3]), k = 6, extra = 1000;
    const int64_t nnz = static_cast<int64_t>(n) * k;
    std::vector<int64_t> cp(n + 1);
    std::vector<int32_t> ri(nnz);
    std::vector<double> v(nnz), c(n), lc(n), uc(n), lr(m), ur(m), x0(n), y0(m);
    if (vj_gen_random_lp(7, m, n, k, 0.0, 0.0, cp.data(), ri.data(), v.data(), c.data(), lc.data(), uc.data(),
                         lr.data(), ur.data(), x0.data(), y0.data()) != 0)
        return 3;
    std::vector<indus::la::Triplet> t;
    t.reserve(static_cast<size_t>(nnz) + extra);
    for (int j = 0; j < n; ++j)
        for (int64_t p = cp[j]; p < cp[j + 1]; ++p) t.push_back({ri[p], j, v[p]});
    indus::Model M;
    M.num_rows = m + 2 * extra;
    M.num_cols = n;
    M.c = c;
    M.col_lower = lc;
    M.col_upper = uc;
    M.row_lower = lr;
    M.row_upper = ur;
    for (int e = 0; e < extra; ++e) {            // empty rows
        M.row_lower.push_back(-1.0);
        M.row_upper.push_back(1.0);
    }
    for (int e = 0; e < extra; ++e) {            // singleton rows x_j <= 20 (looser than the bound 10)
        const int j = (e * 7919) % n;
        t.push_back({m + extra + e, j, 1.0});
        M.row_lower.push_back(-1e20);
        M.row_upper.push_back(20.0);
    }
    for (int j = 0; j < n; j += 50) M.col_lower[j] = M.col_upper[j] = x0[j];   // fixed columns
    M.A.set_from_triplets(M.num_rows, n, t);
    const auto t0 = std::chrono::steady_clock::now();
    const auto pr = indus::presolve::PresolveEngine::apply(M);
    const auto t1 = std::chrono::steady_clock::now();
    const indus::Model& R = pr.presolved_model;
    indus::Solution red;
    red.status = indus::SolveStatus::kOptimal;
    red.col_value.resize(R.num_cols);
    for (int jr = 0; jr < R.num_cols; ++jr) red.col_value[jr] = x0[pr.stack.reduced_to_orig_col[jr]];
    red.row_dual.assign(R.num_rows, 0.0);
    R.A.multiply(red.col_value, red.row_value);
    const auto t2 = std::chrono::steady_clock::now();
    const indus::Solution full = indus::presolve::PresolveEngine::postsolve(red, pr.stack, M);
    const auto t3 = std::chrono::steady_clock::now();
    double maxdiff = 0.0;
    for (int j = 0; j < n; ++j) maxdiff = std::max(maxdiff, std::fabs(full.col_value[j] - x0[j]));
    const auto ver = indus::verifier::verify_solution(M, full, 1e-6);
    std::printf("presolve m=%d n=%d nnz=%lld reductions=%d (fixed %d, empty rows %d, singleton rows %d, "
                "redundant %d) reduced %dx%d apply_s=%.3f postsolve_s=%.3f max|x-x0|=%.3g primal_feasible=%d\n",
                M.num_rows, n, static_cast<long long>(M.A.nnz()), pr.total_reductions(), pr.num_fixed_cols,
                pr.num_empty_rows, pr.num_singleton_rows, pr.num_redundant_rows, R.num_rows, R.num_cols,
                std::chrono::duration<double>(t1 - t0).count(), std::chrono::duration<double>(t3 - t2).count(),
                maxdiff, ver.primal_feasible ? 1 : 0);
    // a singleton row on a fixed column empties first and is removed as an empty row
    const bool ok = pr.num_fixed_cols >= (n + 49) / 50 && pr.num_empty_rows >= extra &&
                    pr.num_empty_rows + pr.num_singleton_rows + pr.num_redundant_rows >= 2 * extra &&
                    R.num_rows <= m && maxdiff == 0.0 && ver.primal_feasible;
    return ok ? 0 : 1;
}

// dump_model <file> <mps|lp>: the Model a reader builds, in hexadecimal floating point, for
// comparison with the reference's own readers.
int dump_model(int argc, char** argv) {
    if (argc < 4) return 2;
    const std::string kind = argv[3];
    const indus::Model M = kind == "lp" ? indus::io::read_lp(argv[2]) : indus::io::read_mps(argv[2]);
    int ints = 0;
    for (auto t : M.col_type) ints += t == indus::VarType::kInteger;
    std::printf("dims %d %d %lld %d %d %lld %a\n", M.num_rows, M.num_cols, static_cast<long long>(M.A.nnz()),
                M.sense == indus::ObjSense::kMaximize ? 1 : 0, ints, static_cast<long long>(M.Q.nnz()),
                M.objective_offset);
    auto ints_line = [](const char* tag, const auto& v) {
        std::printf("%s", tag);
        for (auto a : v) std::printf(" %lld", static_cast<long long>(a));
        std::printf("\n");
    };
    auto dbl_line = [](const char* tag, const std::vector<double>& v) {
        std::printf("%s", tag);
        for (double a : v) std::printf(" %a", a);
        std::printf("\n");
    };
    ints_line("col_ptr", M.A.col_ptr);
    ints_line("row_idx", M.A.row_idx);
    dbl_line("vals", M.A.values);
    dbl_line("c", M.c);
    dbl_line("col_lower", M.col_lower);
    dbl_line("col_upper", M.col_upper);
    dbl_line("row_lower", M.row_lower);
    dbl_line("row_upper", M.row_upper);
    return 0;
}

// presolve_dump <model file> <mps|lp> <out>: the C++ API's presolve and postsolve, written as
// the canonical dump the oracle's shim writes for the reference (tests/test_presolve.py)
int presolve_dump_mode(int argc, char** argv) {
    if (argc < 5) return 2;
    const std::string kind = argv[3];
    indus::Model model = kind == "lp" ? indus::io::read_lp(argv[2]) : indus::io::read_mps(argv[2]);
    auto res = indus::presolve::PresolveEngine::apply(model);
    FILE* f = std::fopen(argv[4], "w");
    if (!f) return 1;
    presolve_dump::dump_result(f, res);
    if (!res.is_infeasible && !res.is_unbounded) {
        auto red = presolve_dump::synthetic_reduced<indus::Solution>(res.presolved_model, indus::SolveStatus::kOptimal);
        presolve_dump::dump_solution(f, indus::presolve::PresolveEngine::postsolve(red, res.stack, model));
    }
    std::fclose(f);
    return 0;
}

// api_solve <m> <n> <fp64|fp32|bf16> <tol> <presolve 0|1> [scaled 0|1]: indus::solve on the
// device, end to end through the C++ API, on the random LP family of the benchmark (6
// nonzeros per column; with `scaled`, 10% of the columns scaled by 10^U[-3,3]) with planted
// presolve reductions (every 50th column fixed at its feasible value, 1000 empty rows, 1000
// singleton rows). Prints the status, the objective, the presolve sizes, the device and host
// times and the verifier. The engine's own time limit is 200 s.
int api_solve(int argc, char** argv) {
    if (argc < 7) return 2;
    const int m = std::atoi(argv[2]), n = std::atoi(argv[3]), k = 6, extra = 1000;
    const std::string prec = argv[4];
    const double tol = std::atof(argv[5]);
    const bool pre = std::atoi(argv[6]) != 0;
    const bool scaled = argc > 7 && std::atoi(argv[7]) != 0;
    const int64_t nnz = static_cast<int64_t>(n) * k;
    std::vector<int64_t> cp(n + 1);
    std::vector<int32_t> ri(nnz);
    std::vector<double> v(nnz), c(n), lc(n), uc(n), lr(m), ur(m), x0(n), y0(m);
    if (vj_gen_random_lp(0x5EED2026ULL + 1, m, n, k, scaled ? 0.1 : 0.0, scaled ? 3.0 : 0.0, cp.data(), ri.data(), v.data(), c.data(), lc.data(),
                         uc.data(), lr.data(), ur.data(), x0.data(), y0.data()) != 0)
        return 3;
    std::vector<indus::la::Triplet> t;
    t.reserve(static_cast<size_t>(nnz) + extra);
    for (int j = 0; j < n; ++j)
        for (int64_t p = cp[j]; p < cp[j + 1]; ++p) t.push_back({ri[p], j, v[p]});
    indus::Model M;
    M.name = "api_scale";
    M.num_rows = m + 2 * extra;
    M.num_cols = n;
    M.c = c;
    M.col_lower = lc;
    M.col_upper = uc;
    M.row_lower = lr;
    M.row_upper = ur;
    for (int e = 0; e < extra; ++e) {
        M.row_lower.push_back(-1.0);
        M.row_upper.push_back(1.0);
    }
    for (int e = 0; e < extra; ++e) {
        t.push_back({m + extra + e, (e * 7919) % n, 1.0});
        M.row_lower.push_back(-1e20);
        M.row_upper.push_back(20.0);
    }
    for (int j = 0; j < n; j += 50) M.col_lower[j] = M.col_upper[j] = x0[j];
    M.A.set_from_triplets(M.num_rows, n, t);
    indus::Options o;
    o.set("algorithm", "pdhg_cuda");
    o.use_gpu = true;
    o.enable_presolve = pre;
    o.set("tolerance", tol);
    o.set("pdhg_precision", prec.c_str());
    o.iteration_limit = 1000000;
    o.time_limit = 200.0;
    const auto t0 = std::chrono::steady_clock::now();
    const indus::Solution s = indus::solve(M, o);
    const double wall = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    const auto g = indus::gpu::get_last_gpu_timing();
    const auto ver = indus::verifier::verify_solution(M, s, 1e-6);
    std::printf("api_solve m=%d n=%d nnz=%lld scaled=%d precision=%s tol=%g presolve=%d status=%d objective=%.17g "
                "reduced=%dx%d reductions=%d iterations=%lld wall_s=%.3f device_total_s=%.3f device_iter_s=%.3f "
                "host_s=%.3f rel_kkt=%.3e verified=%d primal_viol=%.3e dual_viol=%.3e check_interval=%d\n",
                M.num_rows, n, static_cast<long long>(M.A.nnz()), scaled ? 1 : 0, prec.c_str(), tol, pre ? 1 : 0,
                static_cast<int>(s.status), s.objective_value, s.presolve_num_rows, s.presolve_num_cols,
                s.presolve_total_reductions, static_cast<long long>(s.iterations), wall, g.total_time_sec,
                g.iteration_time_sec, wall - g.total_time_sec, g.rel_kkt, ver.passed ? 1 : 0, ver.max_primal_violation,
                ver.max_dual_violation, g.check_interval);
    std::printf("status_message %s\n", s.status_message.c_str());
    return 0;
}

// solve_mps <file.mps> <fp64|fp32> <tol> <presolve 0|1> [out.sol out.json]: indus::solve on the device (the
// reader, presolve, the PDHG engine, postsolve and the quality audit of the C++ API) on one
// model file; prints the status, the objective, the reductions and the relative KKT error.
int solve_mps(int argc, char** argv) {
    if (argc < 6) return 2;
    const indus::Model M = indus::io::read_mps(argv[2]);
    indus::Options o;
    o.set("algorithm", "pdhg_cuda");
    o.use_gpu = true;
    o.enable_presolve = std::atoi(argv[5]) != 0;
    o.set("tolerance", std::atof(argv[4]));
    o.set("pdhg_precision", argv[3]);
    o.iteration_limit = 1000000;
    o.time_limit = 100.0;
    const indus::Solution s = indus::solve(M, o);
    if (argc >= 8) {   // optional: the answer written by this build's .sol and JSON writers
        indus::io::write_solution(s, M, argv[6]);
        indus::io::write_json(s, M, argv[7]);
    }
    const auto k = indus::verifier::relative_kkt(M, s);
    std::printf("solve_mps status=%d objective=%.17g reductions=%d iterations=%lld rel_kkt=%.6e\n",
                static_cast<int>(s.status), s.objective_value, s.presolve_total_reductions,
                static_cast<long long>(s.iterations), k.rel_kkt);
    std::printf("status_message %s\n", s.status_message.c_str());
    return 0;
}

int main(int argc, char** argv) {
    if (argc >= 6 && std::string(argv[1]) == "audit_dump") {
        // audit_dump <mps> <point file> <tol> <out>: the host audits of a given point
        const indus::Model model = indus::io::read_mps(argv[2]);
        audit_dump::Point p;
        if (!audit_dump::read_point(argv[3], p)) return 3;
        FILE* f = std::fopen(argv[5], "w");
        if (!f) return 3;
        audit_dump::run<indus::Solution, indus::Model, indus::SolveStatus>(
            f, model, p, std::stod(argv[4]), [](const indus::Model& m, const indus::Solution& s, double t) {
                return indus::verifier::verify_solution(m, s, t);
            });
        std::fclose(f);
        return 0;
    }
    if (argc >= 6 && std::string(argv[1]) == "write_files") {
        // write_files <mps> <point file> <out.sol> <out.json>: this build's writers on the
        // Solution audit_dump::solution_of builds from the point
        const indus::Model model = indus::io::read_mps(argv[2]);
        audit_dump::Point p;
        if (!audit_dump::read_point(argv[3], p)) return 3;
        const auto s = audit_dump::solution_of<indus::Solution, indus::Model, indus::SolveStatus>(model, p);
        indus::io::write_solution(s, model, argv[4]);
        indus::io::write_json(s, model, argv[5]);
        return 0;
    }
    if (argc >= 2 && std::string(argv[1]) == "options_probe") {
        std::cout << options_probe::run<indus::Options>();
        return 0;
    }
    if (argc >= 2 && std::string(argv[1]) == "solve_mps") return solve_mps(argc, argv);
    if (argc >= 2 && std::string(argv[1]) == "dump") return dump(argc, argv);
    if (argc >= 2 && std::string(argv[1]) == "api_solve") return api_solve(argc, argv);
    if (argc >= 2 && std::string(argv[1]) == "presolve_dump") return presolve_dump_mode(argc, argv);
    if (argc >= 2 && std::string(argv[1]) == "dump_model") return dump_model(argc, argv);
    if (argc >= 2 && std::string(argv[1]) == "presolve_scale") return presolve_scale(argc, argv);
    if (argc < 3 || (std::string(argv[1]) != "cpu" && std::string(argv[1]) != "gpu")) {
        std::cerr << "usage: test_indus_api cpu|gpu <data_dir>\n";
        return 2;
    }
    g_gpu = std::string(argv[1]) == "gpu";
    g_dir = argv[2];
    std::vector<std::pair<const char*, std::function<void()>>> tests = {
        {"two_var", test_two_var},
        {"equality", test_equality},
        {"mixed_senses", test_mixed_senses},
        {"free_variable", test_free_variable},
        {"min_max_symmetry", test_min_max_symmetry},
        {"presolve_handoff", test_presolve_handoff},
        {"status_guard", test_status_guard},
        {"infeasible", test_infeasible},
        {"iteration_limit", test_iteration_limit},
        {"netlib", test_netlib},
        {"device_and_fallback", test_device_and_fallback},
        {"mps_reader", test_mps_reader},
        {"lp_reader", test_lp_reader},
        {"writers", test_writers},
        {"verifier_rejects", test_verifier_rejects},
        {"time_limit_zero", test_time_limit_zero},
    };
    if (g_gpu) {
        tests.push_back({"cpu_gpu_reference_map_identical", test_cpu_gpu_reference_map_identical});
        tests.push_back({"batch", test_batch});
        tests.push_back({"gpu_trivial_sizes", test_gpu_trivial_sizes});
        tests.push_back({"mixed_precision_option", test_mixed_precision_option});
    }
    for (auto& [name, fn] : tests) {
        const int before = g_fail;
        std::cout << "[" << argv[1] << "] " << name << std::endl;
        try {
            fn();
        } catch (const std::exception& e) {
            std::cout << "  EXCEPTION: " << e.what() << "\n";
            ++g_fail;
        }
        std::cout << "  " << (g_fail == before ? "ok" : "FAILED") << std::endl;
    }
    std::cout << (g_fail ? "FAILURES: " : "ALL PASSED") << (g_fail ? std::to_string(g_fail) : "") << std::endl;
    return g_fail ? 1 : 0;
}
