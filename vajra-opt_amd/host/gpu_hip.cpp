// GPU PDHG over the C ABI of include/vajra_pdhg.h (replaces src/solvers/gpu/gpu_context.cpp
// and the host part of src/solvers/gpu/pdhg_cuda.cu of the reference).
#include <chrono>
#include <cmath>
#include <cstring>
#include <iostream>
#include <mutex>
#include <stdexcept>

#include "indus/gpu.hpp"
#include "indus/pdhg.hpp"
#include "indus/tolerances.hpp"
#include "vajra_pdhg.h"

namespace indus::gpu {

namespace {

using clk = std::chrono::high_resolution_clock;
double since(clk::time_point t) { return std::chrono::duration<double>(clk::now() - t).count(); }

std::mutex g_timing_mu;
GpuTimingDiagnostics g_last_timing;

void check(int rc, const char* what) {
    if (rc != VJ_OK) throw std::runtime_error(std::string(what) + " failed (" + std::to_string(rc) + "): " + vj_last_error());
}

struct Ctx {
    vj_ctx* c = nullptr;
    explicit Ctx(int device) { check(vj_ctx_create(device, &c), "vj_ctx_create"); }
    ~Ctx() { vj_ctx_destroy(c); }
    Ctx(const Ctx&) = delete;
    Ctx& operator=(const Ctx&) = delete;
};

vj_params params_from(const Options& o) {
    const std::string mode = o.get_string("pdhg_mode", "restarted");
    vj_params p;
    vj_params_default(&p, mode == "reference" ? VJ_MODE_REFERENCE : VJ_MODE_RESTARTED);
    const std::string prec = o.get_string("pdhg_precision", "fp64");
    p.precision = prec == "fp32" ? VJ_PREC_FP32
                : (prec == "bf16" ? VJ_PREC_BF16 : (prec == "fp16" ? VJ_PREC_FP16 : VJ_PREC_FP64));
    p.max_iterations = o.iteration_limit;
    p.time_limit = o.time_limit;
    p.tolerance = o.get_double("tolerance", 1e-6);
    p.gap_tolerance = o.get_double("gap_tolerance", 1e-6);
    if (o.has("check_frequency")) p.check_frequency = static_cast<int>(o.get_int("check_frequency", p.check_frequency));
    p.enable_scaling = o.enable_scaling ? 1 : 0;
    p.use_graph = o.get_bool("pdhg_graph", true) ? 1 : 0;
    if (o.has("pdhg_refine_every_check"))
        p.refine_every_check = static_cast<int>(o.get_int("pdhg_refine_every_check", 0));
    p.verbose = o.get_bool("verbose", false) ? 1 : 0;
    return p;
}

std::string kernel_mode(const vj_params& p, const vj_result& r) {
    if (p.mode == VJ_MODE_REFERENCE) return "reference_map_exact_order";
    std::string s = "restarted_halpern_";
    s += p.precision == VJ_PREC_FP64 ? "fp64"
         : (p.precision == VJ_PREC_FP32 ? "fp32_refined" : (p.precision == VJ_PREC_FP16 ? "fp16_refined" : "bf16_refined"));
    s += r.col_splits_A > 1 || r.col_splits_At > 1 ? "_xcd_split" : (r.lanes_A == 0 ? "_tile_stream" : (r.lanes_A < 0 ? "_sliced_ell" : "_lane_groups"));
    s += p.use_graph ? "_graph" : "";
    return s;
}

struct HostLp {
    std::vector<int32_t> csc_ri;
};

// The Model's CSC only: the engine builds the CSR of A on the device.
vj_lp to_vj(const Model& M, HostLp& h) {
    h.csc_ri.assign(M.A.row_idx.begin(), M.A.row_idx.end());
    vj_lp lp;
    std::memset(&lp, 0, sizeof(lp));
    lp.m = M.num_rows;
    lp.n = M.num_cols;
    lp.nnz = M.A.nnz();
    lp.csr_row_ptr = nullptr;
    lp.csr_col_idx = nullptr;
    lp.csr_vals = nullptr;
    lp.csc_col_ptr = M.A.col_ptr.data();
    lp.csc_row_idx = h.csc_ri.data();
    lp.csc_vals = M.A.values.data();
    lp.c = M.c.data();
    lp.col_lower = M.col_lower.data();
    lp.col_upper = M.col_upper.data();
    lp.row_lower = M.row_lower.data();
    lp.row_upper = M.row_upper.data();
    lp.maximize = M.sense == ObjSense::kMaximize ? 1 : 0;
    lp.objective_offset = M.objective_offset;
    return lp;
}

SolveStatus to_status(int s) { return static_cast<SolveStatus>(s); }

// Fill a Solution from an unscaled point in the canonical dual convention (pdhg.cpp:380-437
// finalisation: row activity, objective, reduced costs, basis heuristic, quality).
void finish_solution(const Model& M, Solution& sol, const std::vector<double>& x, const std::vector<double>& ycan) {
    const int m = M.num_rows, n = M.num_cols;
    sol.col_value = x;
    sol.row_dual = ycan;
    M.A.multiply(sol.col_value, sol.row_value);
    double obj = M.objective_offset;
    for (int j = 0; j < n; ++j) obj += M.c[j] * x[j];
    sol.objective_value = obj;
    sol.best_dual_bound = obj;
    std::vector<double> aty(n);
    M.A.multiply_transpose(sol.row_dual, aty);
    sol.col_dual.resize(n);
    sol.col_basis_status.assign(n, BasisStatus::kUnknown);
    sol.row_basis_status.assign(m, BasisStatus::kUnknown);
    for (int j = 0; j < n; ++j) {
        sol.col_dual[j] = M.c[j] - aty[j];
        const double lj = M.col_lower[j], uj = M.col_upper[j];
        if (std::fabs(lj - uj) <= tol::kZeroDrop) sol.col_basis_status[j] = BasisStatus::kFixed;
        else if (std::fabs(x[j] - lj) <= 1e-5) sol.col_basis_status[j] = BasisStatus::kAtLower;
        else if (std::fabs(x[j] - uj) <= 1e-5) sol.col_basis_status[j] = BasisStatus::kAtUpper;
        else sol.col_basis_status[j] = BasisStatus::kBasic;
    }
    sol.recompute_quality(M);
}

std::string status_text(const vj_params& p, const vj_result& r) {
    switch (r.status) {
        case VJ_STATUS_OPTIMAL:
            return p.mode == VJ_MODE_REFERENCE ? "PDHG converged to optimal primal and dual tolerance"
                                               : "PDHG reached the relative KKT tolerance";
        case VJ_STATUS_ITERATION_LIMIT:
        case VJ_STATUS_FEASIBLE: return "PDHG reached iteration limit (" + std::to_string(p.max_iterations) + ")";
        case VJ_STATUS_TIME_LIMIT: return "PDHG time limit reached";
        case VJ_STATUS_NUMERICAL_ERROR: return "NaN or Inf detected by the device";
        default: return "";
    }
}

}  // namespace

GpuDeviceInfo probe_device() noexcept {
    GpuDeviceInfo info;
    vj_device_info d;
    if (vj_device_probe(&d) != VJ_OK || !d.available) {
        info.device_name = "None (no HIP device found)";
        return info;
    }
    info.available = true;
    info.device_name = std::string(d.name) + " (" + d.arch + ")";
    info.total_vram_bytes = d.total_bytes;
    info.free_vram_bytes = d.free_bytes;
    info.compute_capability_major = d.major;
    info.compute_capability_minor = d.minor;
    info.multi_processor_count = d.compute_units;
    return info;
}

bool is_gpu_available() noexcept { return probe_device().available; }
bool is_cuda_available() noexcept { return is_gpu_available(); }

GpuTimingDiagnostics get_last_gpu_timing() noexcept {
    std::lock_guard<std::mutex> g(g_timing_mu);
    return g_last_timing;
}

Solution solve_pdhg_gpu(const Model& model, const Options& options) {
    const GpuDeviceInfo info = probe_device();
    if (!info.available) {
        // no silent substitute for the device: the CPU PDHG runs only when the caller asks
        if (!options.get_bool("allow_cpu_fallback", false))
            throw std::runtime_error("solve_pdhg_gpu: no HIP device available (Options \"allow_cpu_fallback\" runs "
                                     "the CPU PDHG instead)");
        std::cout << "[INDUS-GPU] no HIP device (" << info.device_name << "); allow_cpu_fallback set: running the CPU PDHG.\n";
        GpuTimingDiagnostics t;
        t.device_name = "None (CPU fallback)";
        t.kernel_mode = "CPU reference engine";
        {
            std::lock_guard<std::mutex> g(g_timing_mu);
            g_last_timing = t;
        }
        return pdhg::solve_pdhg_cpu(model, options);
    }
    // empty dimensions go to the CPU path, as the reference does (pdhg_cuda.cu:127-134)
    if (model.num_rows == 0 || model.num_cols == 0 || model.A.nnz() == 0) return pdhg::solve_pdhg_cpu(model, options);

    const auto t0 = clk::now();
    Solution sol;
    sol.algorithm_used = "pdhg_cuda";
    const vj_params p = params_from(options);
    HostLp h;
    const vj_lp lp = to_vj(model, h);
    Ctx ctx(0);
    const auto tu = clk::now();
    check(vj_problem_upload(ctx.c, &lp), "vj_problem_upload");
    const double upload = since(tu);
    check(vj_params_set(ctx.c, &p), "vj_params_set");
    vj_result r;
    check(vj_solve(ctx.c, &r), "vj_solve");
    const auto td = clk::now();
    std::vector<double> x(model.num_cols), ycan(model.num_rows);
    check(vj_download(ctx.c, x.data(), nullptr, ycan.data()), "vj_download");
    const double download = since(td);
    sol.status = to_status(r.status);
    sol.status_message = status_text(p, r);
    sol.iterations = r.iterations;
    sol.relative_gap = r.rel_gap;
    finish_solution(model, sol, x, ycan);
    sol.solve_time_seconds = since(t0);

    GpuTimingDiagnostics t;
    t.upload_time_sec = upload;
    t.iteration_time_sec = r.iteration_time;
    t.download_time_sec = download;
    t.total_time_sec = sol.solve_time_seconds;
    t.host_device_transfers = r.host_device_transfers + 2;
    t.iterations = r.iterations;
    t.device_name = info.device_name;
    t.kernel_mode = kernel_mode(p, r);
    t.setup_time_sec = r.setup_time;
    t.rel_kkt = r.rel_kkt;
    t.restarts = r.restart_count;
    t.lanes_A = r.lanes_A;
    t.lanes_At = r.lanes_At;
    t.check_interval = r.check_interval;
    t.col_splits_A = r.col_splits_A;
    t.col_splits_At = r.col_splits_At;
    {
        std::lock_guard<std::mutex> g(g_timing_mu);
        g_last_timing = t;
    }
    if (options.get_bool("verbose", false))
        std::cout << "[INDUS-GPU] " << to_string(sol.status) << " iterations=" << r.iterations << " rel_kkt=" << r.rel_kkt
                  << " mode=" << t.kernel_mode << "\n";
    return sol;
}

std::vector<Solution> solve_pdhg_gpu_batch(const std::vector<Model>& models, const Options& options) {
    std::vector<Solution> out(models.size());
    if (models.empty()) return out;
    if (!is_gpu_available()) {
        if (!options.get_bool("allow_cpu_fallback", false))
            throw std::runtime_error("solve_pdhg_gpu_batch: no HIP device available (Options \"allow_cpu_fallback\" "
                                     "runs the CPU PDHG instead)");
        for (size_t q = 0; q < models.size(); ++q) out[q] = pdhg::solve_pdhg_cpu(models[q], options);
        return out;
    }
    const auto t0 = clk::now();
    const int cnt = static_cast<int>(models.size());
    std::vector<int32_t> ms(cnt), ns(cnt);
    std::vector<int64_t> zs(cnt), tp;
    std::vector<int32_t> ti;
    std::vector<double> tv, c, lc, uc, lr, ur, off(cnt);
    std::vector<int32_t> mx(cnt);
    for (int q = 0; q < cnt; ++q) {
        const Model& M = models[q];
        if (M.num_rows == 0 || M.num_cols == 0 || M.A.nnz() == 0)
            throw std::invalid_argument("solve_pdhg_gpu_batch: every model needs rows, columns and nonzeros");
        ms[q] = M.num_rows;
        ns[q] = M.num_cols;
        zs[q] = M.A.nnz();
        tp.insert(tp.end(), M.A.col_ptr.begin(), M.A.col_ptr.end());
        ti.insert(ti.end(), M.A.row_idx.begin(), M.A.row_idx.end());
        tv.insert(tv.end(), M.A.values.begin(), M.A.values.end());
        c.insert(c.end(), M.c.begin(), M.c.end());
        mx[q] = M.sense == ObjSense::kMaximize ? 1 : 0;
        off[q] = M.objective_offset;
        lc.insert(lc.end(), M.col_lower.begin(), M.col_lower.end());
        uc.insert(uc.end(), M.col_upper.begin(), M.col_upper.end());
        lr.insert(lr.end(), M.row_lower.begin(), M.row_lower.end());
        ur.insert(ur.end(), M.row_upper.begin(), M.row_upper.end());
    }
    // CSCs only: the engine builds the CSRs on the device
    vj_batch b{cnt, ms.data(), ns.data(), zs.data(), nullptr, nullptr, nullptr, tp.data(), ti.data(), tv.data(),
               c.data(), lc.data(), uc.data(), lr.data(), ur.data(), mx.data(), off.data()};
    Ctx ctx(0);
    check(vj_batch_upload(ctx.c, &b), "vj_batch_upload");
    vj_params p = params_from(options);
    p.mode = VJ_MODE_RESTARTED;
    check(vj_params_set(ctx.c, &p), "vj_params_set");
    std::vector<vj_result> res(cnt);
    check(vj_batch_solve(ctx.c, res.data()), "vj_batch_solve");
    int64_t nt = 0, mt = 0;
    for (int q = 0; q < cnt; ++q) {
        nt += ns[q];
        mt += ms[q];
    }
    std::vector<double> x(nt), y(mt);
    check(vj_batch_download(ctx.c, x.data(), y.data()), "vj_batch_download");
    int64_t xo = 0, yo = 0;
    for (int q = 0; q < cnt; ++q) {
        Solution& s = out[q];
        s.algorithm_used = "pdhg_cuda_batch";
        s.status = to_status(res[q].status);
        s.status_message = status_text(p, res[q]);
        s.iterations = res[q].iterations;
        s.relative_gap = res[q].rel_gap;
        finish_solution(models[q], s, std::vector<double>(x.begin() + xo, x.begin() + xo + ns[q]),
                        std::vector<double>(y.begin() + yo, y.begin() + yo + ms[q]));
        s.solve_time_seconds = since(t0);
        xo += ns[q];
        yo += ms[q];
    }
    return out;
}

}  // namespace indus::gpu
