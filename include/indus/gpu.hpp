// GPU PDHG entry points (restates include/indus/gpu.hpp:16-59 of the reference), backed by
// the MI355X engine through the C ABI of include/vajra_pdhg.h.
#pragma once
#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "indus/model.hpp"
#include "indus/options.hpp"

namespace indus::gpu {

using indus::Model;
using indus::Options;
using indus::Solution;

struct GpuDeviceInfo {
    bool available = false;
    std::string device_name = "None";
    size_t total_vram_bytes = 0;
    size_t free_vram_bytes = 0;
    int compute_capability_major = 0;   // gfx major / minor on AMD devices
    int compute_capability_minor = 0;
    int multi_processor_count = 0;      // compute units
};

// Kept for source compatibility (gpu.hpp:27-34); the engine's per-check record is richer.
struct DeviceDiagnostics {
    double primal_residual = 0.0;
    double dual_residual = 0.0;
    double objective = 0.0;
    double duality_gap = 0.0;
    int converged = 0;
    int numerical_error = 0;
};

struct GpuTimingDiagnostics {
    double upload_time_sec = 0.0;
    double iteration_time_sec = 0.0;
    double download_time_sec = 0.0;
    double total_time_sec = 0.0;
    int64_t host_device_transfers = 0;
    int64_t iterations = 0;
    std::string device_name = "None";
    std::string kernel_mode = "none";
    // engine extras
    double setup_time_sec = 0.0;
    double rel_kkt = 0.0;
    int restarts = 0;
    int lanes_A = 0, lanes_At = 0, check_interval = 0, col_splits_A = 1, col_splits_At = 1;
};

[[nodiscard]] GpuDeviceInfo probe_device() noexcept;
// Name kept from the reference for drop-in use: true when a HIP device is available.
[[nodiscard]] bool is_cuda_available() noexcept;
[[nodiscard]] bool is_gpu_available() noexcept;
[[nodiscard]] GpuTimingDiagnostics get_last_gpu_timing() noexcept;

// GPU PDHG. With no device it throws std::runtime_error, unless Options "allow_cpu_fallback"
// is true: then it prints a notice and runs the CPU PDHG, as the reference always does
// (gpu_context.cpp:65-71). HIP failures throw std::runtime_error (pdhg_cuda.cu:20-28).
Solution solve_pdhg_gpu(const Model& model, const Options& options = Options());

// Many independent LPs in one device pass (no reference counterpart): every instance keeps
// its own restart and convergence state; results come back in input order.
std::vector<Solution> solve_pdhg_gpu_batch(const std::vector<Model>& models, const Options& options = Options());

}  // namespace indus::gpu
