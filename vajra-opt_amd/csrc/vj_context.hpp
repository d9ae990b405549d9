// Context object behind the opaque vj_ctx handle.
#pragma once
#include <memory>
#include <string>
#include <vector>

#include "vj_internal.hpp"

namespace vj {

struct RefEngine;   // reference iteration map (vj_reference.hip)
struct FastEngine;  // restarted Halpern engine, single LP and batches (vj_engine.hip)
void destroy_ref(RefEngine* e);
void destroy_fast(FastEngine* e);
struct RefDeleter { void operator()(RefEngine* e) const { destroy_ref(e); } };
struct FastDeleter { void operator()(FastEngine* e) const { destroy_fast(e); } };

// The problem as uploaded: original (unscaled) data, A in CSR and CSR(A^T), possibly a
// block-diagonal pack of `count` instances with global indices.
struct Problem {
    int count = 0;
    int m = 0, n = 0;                      // totals
    int64_t nnz = 0;
    std::vector<int32_t> m_i, n_i;         // per instance
    std::vector<int64_t> nnz_i;
    std::vector<int64_t> row_off, col_off; // count+1 prefix offsets
    std::vector<double> offset_i;          // objective offsets (canonical sign applied)
    std::vector<int32_t> maximize_i;
    // host copies (global indices, int64 pointers)
    std::vector<int64_t> h_rp, h_tp;
    std::vector<int32_t> h_ci, h_ti;
    std::vector<double> h_av, h_tv;
    std::vector<double> h_c, h_lc, h_uc, h_lr, h_ur;  // as given (c not negated)
    // device copies of the original problem (int32 pointers)
    DevBuf rp, ci, av, tp, ti, tv;
    DevBuf c_can, lc, uc, lr, ur;          // canonical-min c; bounds with +-inf
    DevBuf row_inst, col_inst;             // instance id per row / column (count > 1)
    DevBuf d_row_off, d_col_off;           // int32 count+1
    std::vector<double> norm_b, norm_c;    // original-space norms per instance
    bool all_eq = false;                   // every row of every instance has l = u (finite)

    // Back to the empty state, releasing device memory (DevBuf is not copyable).
    void reset() {
        count = 0;
        m = n = 0;
        nnz = 0;
        for (auto* v : {&m_i, &n_i, &maximize_i}) v->clear();
        for (auto* v : {&nnz_i, &row_off, &col_off, &h_rp, &h_tp}) v->clear();
        for (auto* v : {&offset_i, &h_av, &h_tv, &h_c, &h_lc, &h_uc, &h_lr, &h_ur, &norm_b, &norm_c}) v->clear();
        all_eq = false;
        h_ci.clear();
        h_ti.clear();
        for (DevBuf* b : {&rp, &ci, &av, &tp, &ti, &tv, &c_can, &lc, &uc, &lr, &ur, &row_inst, &col_inst,
                          &d_row_off, &d_col_off})
            b->release();
    }
};

struct KernelTimer {
    std::string name;
    double total_ms = 0.0;
    int64_t launches = 0;
    double bytes = 0.0;
};

}  // namespace vj

struct vj_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    vj_params params{};
    bool params_given = false;
    bool uploaded = false;
    bool batch = false;
    vj::Problem prob;
    // returned point in the original space: x and the canonical-min dual
    vj::DevBuf x_sol, y_sol;
    bool have_solution = false;
    std::vector<double> ref_row_dual;      // reference-convention dual of the last REFERENCE solve
    vj_result last{};
    std::vector<vj_result> batch_last;
    std::vector<vj::KernelTimer> timers;
    std::unique_ptr<vj::RefEngine, vj::RefDeleter> ref;
    std::unique_ptr<vj::FastEngine, vj::FastDeleter> fast;
    ~vj_ctx();
};

namespace vj {
// internal entry points implemented across translation units
int reference_solve(vj_ctx* ctx, vj_result* out, int64_t stop_at_iter, double* traj_x,
                    double* traj_xbar, double* traj_y);
int fast_solve(vj_ctx* ctx, vj_result* out_array);
int fast_benchmark(vj_ctx* ctx, int reps, vj_kernel_stat* out, int cap, int* written);
int fast_scaling(vj_ctx* ctx, double* row_scale, double* col_scale);
int fast_benchmark_blocks(vj_ctx* ctx, int warmup, int steps, double* ms_per_block, int* iters_per_block);
// FP64 rel-KKT audit of device vectors (original space, canonical dual) for all
// instances; out8 has count*8 entries.
int audit_device(vj_ctx* ctx, const double* d_x, const double* d_y, double* out8);
}  // namespace vj
