// C-ABI entry points of libvajra_hip.so (declared in include/vajra_pdhg.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <limits>
#include <new>
#include <string>

#include "vj_context.hpp"
#include "vj_layout.hpp"

namespace vj {
namespace {
thread_local std::string g_last_error;
}
void set_error(const std::string& msg) { g_last_error = msg; }
// Entry points start from a clean slate: the HIP runtime's sticky last-error may hold a
// code left by another library in the process (e.g. a framework's own device probing).
void clear_error() {
    g_last_error.clear();
    (void)hipGetLastError();
}
}  // namespace vj

vj_ctx::~vj_ctx() {
    ref.reset();
    fast.reset();
    if (stream) (void)hipStreamDestroy(stream);
}

using namespace vj;

namespace {

// Device data common to all modes: int32 pointers, original values, canonical c,
// bounds with +-inf for infinite entries, instance ids and offsets.
int upload_common(vj_ctx* ctx) {
    Problem& P = ctx->prob;
    hipStream_t s = ctx->stream;
    if (P.nnz >= static_cast<int64_t>(std::numeric_limits<int32_t>::max())) {
        set_error("nnz must be below 2^31");
        return VJ_ERR_INVALID_ARGUMENT;
    }
    const bool csr_given = !P.h_rp.empty();
    std::vector<int32_t> rp(csr_given ? P.m + 1 : 0), tp(P.n + 1);
    for (size_t i = 0; i < rp.size(); ++i) rp[i] = static_cast<int32_t>(P.h_rp[i]);
    for (int j = 0; j <= P.n; ++j) tp[j] = static_cast<int32_t>(P.h_tp[j]);
    std::vector<double> ccan(P.n), lc(P.n), uc(P.n), lr(P.m), ur(P.m);
    std::vector<int32_t> rinst(P.m), cinst(P.n), roff(P.count + 1), coff(P.count + 1);
    const double inf = std::numeric_limits<double>::infinity();
    auto fix_lo = [&](double v) { return v <= -1e20 ? -inf : v; };
    auto fix_hi = [&](double v) { return v >= 1e20 ? inf : v; };
    for (int q = 0; q < P.count; ++q) {
        const double sense = P.maximize_i[q] ? -1.0 : 1.0;
        for (int64_t j = P.col_off[q]; j < P.col_off[q + 1]; ++j) {
            ccan[j] = sense * P.h_c[j];
            cinst[j] = q;
        }
        for (int64_t i = P.row_off[q]; i < P.row_off[q + 1]; ++i) rinst[i] = q;
        roff[q] = static_cast<int32_t>(P.row_off[q]);
        coff[q] = static_cast<int32_t>(P.col_off[q]);
    }
    roff[P.count] = P.m;
    coff[P.count] = P.n;
    for (int j = 0; j < P.n; ++j) {
        lc[j] = fix_lo(P.h_lc[j]);
        uc[j] = fix_hi(P.h_uc[j]);
    }
    P.all_eq = P.m > 0;
    for (int i = 0; i < P.m; ++i) {
        lr[i] = fix_lo(P.h_lr[i]);
        ur[i] = fix_hi(P.h_ur[i]);
        P.all_eq = P.all_eq && lr[i] == ur[i] && std::isfinite(lr[i]);
    }
    // original-space norms for the rel-KKT denominators (finite-bound vector and c)
    P.norm_b.assign(P.count, 0.0);
    P.norm_c.assign(P.count, 0.0);
    for (int q = 0; q < P.count; ++q) {
        double b2 = 0.0, c2 = 0.0;
        for (int64_t i = P.row_off[q]; i < P.row_off[q + 1]; ++i) {
            double b = 0.0;
            if (std::isfinite(lr[i])) b = std::fabs(lr[i]);
            if (std::isfinite(ur[i]) && std::fabs(ur[i]) > b) b = std::fabs(ur[i]);
            b2 += b * b;
        }
        for (int64_t j = P.col_off[q]; j < P.col_off[q + 1]; ++j) c2 += ccan[j] * ccan[j];
        P.norm_b[q] = std::sqrt(b2);
        P.norm_c[q] = std::sqrt(c2);
    }
    auto up = [&](DevBuf& b, const void* h, size_t bytes) -> int {
        VJ_TRY(b.alloc(bytes));
        VJ_HIP(hipMemcpyAsync(b.p, h, bytes, hipMemcpyHostToDevice, s));
        return VJ_OK;
    };
    VJ_TRY(up(P.tp, tp.data(), tp.size() * 4));
    VJ_TRY(up(P.ti, P.h_ti.data(), P.nnz * 4));
    VJ_TRY(up(P.tv, P.h_tv.data(), P.nnz * 8));
    if (csr_given) {
        VJ_TRY(up(P.rp, rp.data(), rp.size() * 4));
        VJ_TRY(up(P.ci, P.h_ci.data(), P.nnz * 4));
        VJ_TRY(up(P.av, P.h_av.data(), P.nnz * 8));
    } else {
        // the CSR of A, transposed on the device from the uploaded CSC (to_csr order)
        VJ_TRY(device_transpose(s, P.n, P.m, P.nnz, P.tp.as<int>(), P.ti.as<int>(), P.tv.as<double>(), P.rp, P.ci,
                                P.av));
    }
    VJ_TRY(up(P.c_can, ccan.data(), P.n * 8));
    VJ_TRY(up(P.lc, lc.data(), P.n * 8));
    VJ_TRY(up(P.uc, uc.data(), P.n * 8));
    VJ_TRY(up(P.lr, lr.data(), P.m * 8));
    VJ_TRY(up(P.ur, ur.data(), P.m * 8));
    VJ_TRY(up(P.row_inst, rinst.data(), P.m * 4));
    VJ_TRY(up(P.col_inst, cinst.data(), P.n * 4));
    VJ_TRY(up(P.d_row_off, roff.data(), roff.size() * 4));
    VJ_TRY(up(P.d_col_off, coff.data(), coff.size() * 4));
    VJ_HIP(hipStreamSynchronize(s));
    return VJ_OK;
}

bool check_ptr_array(const int64_t* p, int len, int64_t nnz) {
    if (p[0] != 0 || p[len] != nnz) return false;
    for (int i = 0; i < len; ++i)
        if (p[i + 1] < p[i]) return false;
    return true;
}
bool check_idx(const int64_t* p, const int32_t* idx, int len, int bound) {
    for (int r = 0; r < len; ++r)
        for (int64_t k = p[r]; k < p[r + 1]; ++k) {
            if (idx[k] < 0 || idx[k] >= bound) return false;
            if (k > p[r] && idx[k] <= idx[k - 1]) return false;
        }
    return true;
}

}  // namespace

extern "C" {

const char* vj_last_error(void) { return g_last_error.c_str(); }

int vj_device_probe(vj_device_info* out) {
    if (!out) return VJ_ERR_INVALID_ARGUMENT;
    std::memset(out, 0, sizeof(*out));
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count <= 0) {
        std::snprintf(out->name, sizeof(out->name), "None (no HIP device)");
        return VJ_OK;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, 0) != hipSuccess) {
        std::snprintf(out->name, sizeof(out->name), "Error querying HIP device");
        return VJ_OK;
    }
    out->available = 1;
    out->device_count = count;
    // some driver stacks leave the marketing name empty; fall back to the architecture
    if (prop.name[0] != '\0') std::snprintf(out->name, sizeof(out->name), "%s", prop.name);
    else std::snprintf(out->name, sizeof(out->name), "AMD Instinct GPU (%d CUs)", prop.multiProcessorCount);
    std::snprintf(out->arch, sizeof(out->arch), "%s", prop.gcnArchName);
    out->compute_units = prop.multiProcessorCount;
    out->wavefront_size = prop.warpSize;
    out->major = prop.major;
    out->minor = prop.minor;
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) == hipSuccess) {
        out->free_bytes = fr;
        out->total_bytes = tot;
    } else {
        out->total_bytes = prop.totalGlobalMem;
    }
    return VJ_OK;
}

void vj_params_default(vj_params* p, int mode) {
    std::memset(p, 0, sizeof(*p));
    p->mode = mode;
    p->precision = VJ_PREC_FP64;
    p->max_iterations = 1000000;
    p->time_limit = 1e20;
    p->tolerance = 1e-6;
    p->gap_tolerance = 1e-6;
    p->check_frequency = mode == VJ_MODE_REFERENCE ? 100 : 0;
    p->restart_stagnation_limit = 200;
    p->step_size_safety = 0.95;
    p->enable_scaling = 1;
    p->ruiz_iterations = 10;
    p->pock_chambolle = 1;
    p->bound_objective_rescaling = 1;
    p->reflection = 1.0;
    p->step_scale = 0.95;
    p->use_graph = 1;
    p->lanes_A = 0;
    p->lanes_At = 0;
    p->refine_rounds = 0;
    p->verbose = 0;
    p->restart_sufficient = 0.2;
    p->restart_necessary = 0.8;
    p->restart_artificial = 0.5;
    p->primal_weight_smoothing = 0.5;
    p->primal_weight_integral = 0.0;
    p->power_iterations = 64;
    p->restart_drift = 0.0;
    p->refine_every_check = 0;
    p->stage_threads = 0;
}

int vj_ctx_create(int device, vj_ctx** out) {
    clear_error();
    if (!out) return VJ_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) {
        set_error("no HIP device available");
        return VJ_ERR_NO_DEVICE;
    }
    if (device < 0 || device >= count) {
        set_error("device index out of range");
        return VJ_ERR_INVALID_ARGUMENT;
    }
    VJ_HIP(hipSetDevice(device));
    vj_ctx* c = new (std::nothrow) vj_ctx();
    if (!c) return VJ_ERR_ALLOC;
    c->device = device;
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        set_error(std::string("hipStreamCreate failed: ") + hipGetErrorString(e));
        delete c;
        return VJ_ERR_LAUNCH;
    }
    vj_params_default(&c->params, VJ_MODE_RESTARTED);
    *out = c;
    return VJ_OK;
}

int vj_ctx_destroy(vj_ctx* ctx) {
    if (!ctx) return VJ_OK;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    delete ctx;
    return VJ_OK;
}

int vj_params_set(vj_ctx* ctx, const vj_params* p) {
    clear_error();
    if (!ctx || !p) return VJ_ERR_INVALID_ARGUMENT;
    if (p->mode != VJ_MODE_REFERENCE && p->mode != VJ_MODE_RESTARTED) {
        set_error("unknown mode");
        return VJ_ERR_INVALID_ARGUMENT;
    }
    if (p->precision < VJ_PREC_FP64 || p->precision > VJ_PREC_FP16) {
        set_error("unknown precision");
        return VJ_ERR_INVALID_ARGUMENT;
    }
    auto lanes_ok = [](int g) { return (g >= -7 && g <= 0) || g == 2 || g == 4 || g == 8 || g == 16 || g == 32; };
    auto splits_ok = [](int k) { return k == 0 || k == 1 || k == 2 || k == 4 || k == 8; };
    if (!splits_ok(p->col_splits_A) || !splits_ok(p->col_splits_At)) {
        set_error("column splits must be 0 (auto), 1, 2, 4 or 8");
        return VJ_ERR_INVALID_ARGUMENT;
    }
    if (!(p->restart_sufficient > 0.0 && p->restart_sufficient < 1.0) ||
        !(p->restart_necessary >= p->restart_sufficient && p->restart_necessary < 1.0) ||
        !(p->restart_artificial > 0.0 && p->restart_artificial <= 1.0) ||
        !(p->primal_weight_smoothing >= 0.0 && p->primal_weight_smoothing <= 1.0) ||
        !(p->primal_weight_integral >= 0.0 && p->primal_weight_integral < 1.0) || p->power_iterations < 1 ||
        p->power_iterations > 10000 || std::isnan(p->restart_drift) ||
        (p->restart_drift > 0.0 && p->restart_drift <= 1.0) || p->refine_every_check < -1 ||
        p->refine_every_check > 1 || (p->stage_threads != 0 && p->stage_threads != 256 && p->stage_threads != 1024)) {
        set_error("restart, primal-weight, power-iteration, refine_every_check or stage_threads parameter out of range");
        return VJ_ERR_INVALID_ARGUMENT;
    }
    if (!lanes_ok(p->lanes_A) || !lanes_ok(p->lanes_At)) {
        set_error("lanes per row must be 0 (autotune), -1 (tile-streamed), -2 (tile-streamed, FP32 records), -3 "
                  "(sliced ELL), -4 (sliced ELL, FP32 records), -5 (LDS-staged), -6 (LDS-staged, FP32 records), -7 "
                  "(LDS-staged, 16-bit window offsets) or one of 2, 4, 8, 16, 32");
        return VJ_ERR_INVALID_ARGUMENT;
    }
    if (p->max_iterations < 0 || p->check_frequency < 0 || !(p->tolerance >= 0.0) ||
        !(p->step_scale > 0.0 && p->step_scale < 1.0) || p->reflection < 0.0 || p->reflection > 1.0) {
        set_error("invalid parameter value");
        return VJ_ERR_INVALID_ARGUMENT;
    }
    ctx->params = *p;
    ctx->params_given = true;
    return VJ_OK;
}

int vj_problem_upload(vj_ctx* ctx, const vj_lp* lp) {
    clear_error();
    if (!ctx || !lp) return VJ_ERR_INVALID_ARGUMENT;
    if (lp->m <= 0 || lp->n <= 0 || lp->nnz <= 0) {
        set_error("the device engine needs m > 0, n > 0 and nnz > 0 (callers solve empty LPs on the host)");
        return VJ_ERR_INVALID_ARGUMENT;
    }
    const int csr_parts = (lp->csr_row_ptr ? 1 : 0) + (lp->csr_col_idx ? 1 : 0) + (lp->csr_vals ? 1 : 0);
    if ((csr_parts != 0 && csr_parts != 3) || !lp->csc_col_ptr || !lp->csc_row_idx || !lp->csc_vals || !lp->c ||
        !lp->col_lower || !lp->col_upper || !lp->row_lower || !lp->row_upper) {
        set_error("null array in vj_lp (the CSR arrays may be NULL together: the CSR is then built on the device)");
        return VJ_ERR_INVALID_ARGUMENT;
    }
    const bool csr = csr_parts == 3;
    if (!check_ptr_array(lp->csc_col_ptr, lp->n, lp->nnz) || !check_idx(lp->csc_col_ptr, lp->csc_row_idx, lp->n, lp->m) ||
        (csr && (!check_ptr_array(lp->csr_row_ptr, lp->m, lp->nnz) ||
                 !check_idx(lp->csr_row_ptr, lp->csr_col_idx, lp->m, lp->n)))) {
        set_error("malformed CSR/CSC arrays (pointers must be monotone from 0 to nnz, indices in range and strictly ascending per row)");
        return VJ_ERR_INVALID_ARGUMENT;
    }
    VJ_HIP(hipSetDevice(ctx->device));
    Problem& P = ctx->prob;
    P.reset();
    P.count = 1;
    P.m = lp->m;
    P.n = lp->n;
    P.nnz = lp->nnz;
    P.m_i = {lp->m};
    P.n_i = {lp->n};
    P.nnz_i = {lp->nnz};
    P.row_off = {0, lp->m};
    P.col_off = {0, lp->n};
    P.maximize_i = {lp->maximize ? 1 : 0};
    P.offset_i = {(lp->maximize ? -1.0 : 1.0) * lp->objective_offset};
    if (csr) {
        P.h_rp.assign(lp->csr_row_ptr, lp->csr_row_ptr + lp->m + 1);
        P.h_ci.assign(lp->csr_col_idx, lp->csr_col_idx + lp->nnz);
        P.h_av.assign(lp->csr_vals, lp->csr_vals + lp->nnz);
    }
    P.h_tp.assign(lp->csc_col_ptr, lp->csc_col_ptr + lp->n + 1);
    P.h_ti.assign(lp->csc_row_idx, lp->csc_row_idx + lp->nnz);
    P.h_tv.assign(lp->csc_vals, lp->csc_vals + lp->nnz);
    P.h_c.assign(lp->c, lp->c + lp->n);
    P.h_lc.assign(lp->col_lower, lp->col_lower + lp->n);
    P.h_uc.assign(lp->col_upper, lp->col_upper + lp->n);
    P.h_lr.assign(lp->row_lower, lp->row_lower + lp->m);
    P.h_ur.assign(lp->row_upper, lp->row_upper + lp->m);
    ctx->fast.reset();
    ctx->ref.reset();
    ctx->have_solution = false;
    ctx->batch = false;
    const auto t0 = std::chrono::steady_clock::now();
    VJ_TRY(upload_common(ctx));
    ctx->last = vj_result{};
    ctx->last.upload_time = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    ctx->uploaded = true;
    return VJ_OK;
}

int vj_batch_upload(vj_ctx* ctx, const vj_batch* b) {
    clear_error();
    if (!ctx || !b || b->count <= 0) return VJ_ERR_INVALID_ARGUMENT;
    const int csr_parts = (b->csr_row_ptr ? 1 : 0) + (b->csr_col_idx ? 1 : 0) + (b->csr_vals ? 1 : 0);
    if (!b->m || !b->n || !b->nnz || (csr_parts != 0 && csr_parts != 3) || !b->csc_col_ptr || !b->csc_row_idx ||
        !b->csc_vals || !b->c || !b->col_lower || !b->col_upper || !b->row_lower || !b->row_upper) {
        set_error("null array in vj_batch (the CSR arrays may be NULL together: the CSR is then built on the device)");
        return VJ_ERR_INVALID_ARGUMENT;
    }
    const bool csr = csr_parts == 3;
    VJ_HIP(hipSetDevice(ctx->device));
    Problem& P = ctx->prob;
    P.reset();
    P.count = b->count;
    P.row_off.assign(b->count + 1, 0);
    P.col_off.assign(b->count + 1, 0);
    std::vector<int64_t> nnz_off(b->count + 1, 0);
    for (int q = 0; q < b->count; ++q) {
        if (b->m[q] <= 0 || b->n[q] <= 0 || b->nnz[q] <= 0) {
            set_error("every batch instance needs m > 0, n > 0 and nnz > 0");
            return VJ_ERR_INVALID_ARGUMENT;
        }
        P.row_off[q + 1] = P.row_off[q] + b->m[q];
        P.col_off[q + 1] = P.col_off[q] + b->n[q];
        nnz_off[q + 1] = nnz_off[q] + b->nnz[q];
    }
    if (P.row_off[b->count] >= std::numeric_limits<int32_t>::max() ||
        P.col_off[b->count] >= std::numeric_limits<int32_t>::max()) {
        set_error("batch too large for 32-bit row/column indices");
        return VJ_ERR_INVALID_ARGUMENT;
    }
    P.m = static_cast<int>(P.row_off[b->count]);
    P.n = static_cast<int>(P.col_off[b->count]);
    P.nnz = nnz_off[b->count];
    P.m_i.assign(b->m, b->m + b->count);
    P.n_i.assign(b->n, b->n + b->count);
    P.nnz_i.assign(b->nnz, b->nnz + b->count);
    P.maximize_i.assign(b->count, 0);
    P.offset_i.assign(b->count, 0.0);
    for (int q = 0; q < b->count; ++q) {
        const bool mx = b->maximize && b->maximize[q] != 0;
        const double off = b->objective_offset ? b->objective_offset[q] : 0.0;
        if (!std::isfinite(off)) {
            set_error("non-finite objective offset in batch instance " + std::to_string(q));
            return VJ_ERR_INVALID_ARGUMENT;
        }
        P.maximize_i[q] = mx ? 1 : 0;
        P.offset_i[q] = (mx ? -1.0 : 1.0) * off;   // canonical (minimisation) sign, as for one LP
    }
    if (csr) {
        P.h_rp.resize(P.m + 1);
        P.h_ci.resize(P.nnz);
        P.h_av.assign(b->csr_vals, b->csr_vals + P.nnz);
    }
    P.h_tp.resize(P.n + 1);
    P.h_ti.resize(P.nnz);
    P.h_tv.assign(b->csc_vals, b->csc_vals + P.nnz);
    for (int q = 0; q < b->count; ++q) {
        const int64_t r0 = P.row_off[q], c0 = P.col_off[q], z0 = nnz_off[q];
        const int64_t* rp = csr ? b->csr_row_ptr + r0 + q : nullptr;  // each instance has m+1 pointers
        const int64_t* tp = b->csc_col_ptr + c0 + q;
        if (!check_ptr_array(tp, b->n[q], b->nnz[q]) || !check_idx(tp, b->csc_row_idx + z0, b->n[q], b->m[q]) ||
            (csr && (!check_ptr_array(rp, b->m[q], b->nnz[q]) || !check_idx(rp, b->csr_col_idx + z0, b->m[q], b->n[q])))) {
            set_error("malformed CSR/CSC arrays in batch instance " + std::to_string(q));
            return VJ_ERR_INVALID_ARGUMENT;
        }
        for (int j = 0; j < b->n[q]; ++j) P.h_tp[c0 + j] = z0 + tp[j];
        for (int64_t k = 0; k < b->nnz[q]; ++k) P.h_ti[z0 + k] = static_cast<int32_t>(r0 + b->csc_row_idx[z0 + k]);
        if (csr) {
            for (int i = 0; i < b->m[q]; ++i) P.h_rp[r0 + i] = z0 + rp[i];
            for (int64_t k = 0; k < b->nnz[q]; ++k) P.h_ci[z0 + k] = static_cast<int32_t>(c0 + b->csr_col_idx[z0 + k]);
        }
    }
    if (csr) P.h_rp[P.m] = P.nnz;
    P.h_tp[P.n] = P.nnz;
    P.h_c.assign(b->c, b->c + P.n);
    P.h_lc.assign(b->col_lower, b->col_lower + P.n);
    P.h_uc.assign(b->col_upper, b->col_upper + P.n);
    P.h_lr.assign(b->row_lower, b->row_lower + P.m);
    P.h_ur.assign(b->row_upper, b->row_upper + P.m);
    ctx->fast.reset();
    ctx->ref.reset();
    ctx->have_solution = false;
    ctx->batch = true;
    VJ_TRY(upload_common(ctx));
    ctx->uploaded = true;
    return VJ_OK;
}

int vj_solve(vj_ctx* ctx, vj_result* out) {
    clear_error();
    if (!ctx || !out) return VJ_ERR_INVALID_ARGUMENT;
    if (!ctx->uploaded || ctx->batch) {
        set_error(ctx->batch ? "a batch is uploaded: use vj_batch_solve" : "no problem uploaded");
        return VJ_ERR_NOT_READY;
    }
    VJ_HIP(hipSetDevice(ctx->device));
    const double upload_time = ctx->last.upload_time;
    int rc;
    if (ctx->params.mode == VJ_MODE_REFERENCE) rc = reference_solve(ctx, out, 0, nullptr, nullptr, nullptr);
    else rc = fast_solve(ctx, out);
    if (rc != VJ_OK) return rc;
    if (ctx->params.mode != VJ_MODE_REFERENCE) out->upload_time = upload_time;
    ctx->last = *out;
    return VJ_OK;
}

int vj_batch_solve(vj_ctx* ctx, vj_result* out) {
    clear_error();
    if (!ctx || !out) return VJ_ERR_INVALID_ARGUMENT;
    if (!ctx->uploaded || !ctx->batch) {
        set_error("no batch uploaded");
        return VJ_ERR_NOT_READY;
    }
    if (ctx->params.mode != VJ_MODE_RESTARTED) {
        set_error("batches run in VJ_MODE_RESTARTED");
        return VJ_ERR_UNSUPPORTED;
    }
    VJ_HIP(hipSetDevice(ctx->device));
    ctx->batch_last.resize(ctx->prob.count);
    VJ_TRY(fast_solve(ctx, ctx->batch_last.data()));
    std::memcpy(out, ctx->batch_last.data(), sizeof(vj_result) * ctx->prob.count);
    return VJ_OK;
}

int vj_download(vj_ctx* ctx, double* x, double* row_dual, double* y_canonical) {
    clear_error();
    if (!ctx) return VJ_ERR_INVALID_ARGUMENT;
    if (!ctx->have_solution || ctx->batch) {
        set_error("no single-LP solution available");
        return VJ_ERR_NOT_READY;
    }
    VJ_HIP(hipSetDevice(ctx->device));
    const int m = ctx->prob.m, n = ctx->prob.n;
    if (x) VJ_HIP(hipMemcpy(x, ctx->x_sol.p, static_cast<size_t>(n) * 8, hipMemcpyDeviceToHost));
    std::vector<double> ycan(m);
    VJ_HIP(hipMemcpy(ycan.data(), ctx->y_sol.p, static_cast<size_t>(m) * 8, hipMemcpyDeviceToHost));
    if (y_canonical) std::memcpy(y_canonical, ycan.data(), static_cast<size_t>(m) * 8);
    if (row_dual) {
        // reference convention (pdhg.cpp:409-411): minimise -> y_sol = canonical dual,
        // maximise -> the PDHG dual of the negated problem, i.e. minus the canonical dual
        const bool is_max = ctx->prob.maximize_i[0] != 0;
        for (int i = 0; i < m; ++i) row_dual[i] = is_max ? -ycan[i] : ycan[i];
    }
    return VJ_OK;
}

int vj_batch_download(vj_ctx* ctx, double* x, double* y_canonical) {
    clear_error();
    if (!ctx) return VJ_ERR_INVALID_ARGUMENT;
    if (!ctx->have_solution || !ctx->batch) {
        set_error("no batch solution available");
        return VJ_ERR_NOT_READY;
    }
    VJ_HIP(hipSetDevice(ctx->device));
    if (x) VJ_HIP(hipMemcpy(x, ctx->x_sol.p, static_cast<size_t>(ctx->prob.n) * 8, hipMemcpyDeviceToHost));
    if (y_canonical) VJ_HIP(hipMemcpy(y_canonical, ctx->y_sol.p, static_cast<size_t>(ctx->prob.m) * 8, hipMemcpyDeviceToHost));
    return VJ_OK;
}

int vj_reference_iterates(vj_ctx* ctx, int64_t k, double* x, double* x_bar, double* y) {
    clear_error();
    if (!ctx || k <= 0) return VJ_ERR_INVALID_ARGUMENT;
    if (!ctx->uploaded || ctx->batch) {
        set_error("no single LP uploaded");
        return VJ_ERR_NOT_READY;
    }
    if (ctx->params.mode != VJ_MODE_REFERENCE) {
        set_error("trajectory probes need VJ_MODE_REFERENCE");
        return VJ_ERR_UNSUPPORTED;
    }
    VJ_HIP(hipSetDevice(ctx->device));
    vj_result r;
    return reference_solve(ctx, &r, k, x, x_bar, y);
}

int vj_audit_point(vj_ctx* ctx, const double* x, const double* y_canonical, double* out8) {
    clear_error();
    if (!ctx || !x || !y_canonical || !out8) return VJ_ERR_INVALID_ARGUMENT;
    if (!ctx->uploaded || ctx->batch) {
        set_error("no single LP uploaded");
        return VJ_ERR_NOT_READY;
    }
    VJ_HIP(hipSetDevice(ctx->device));
    DevBuf dx, dy;
    VJ_TRY(dx.alloc(static_cast<size_t>(ctx->prob.n) * 8));
    VJ_TRY(dy.alloc(static_cast<size_t>(ctx->prob.m) * 8));
    VJ_HIP(hipMemcpy(dx.p, x, static_cast<size_t>(ctx->prob.n) * 8, hipMemcpyHostToDevice));
    VJ_HIP(hipMemcpy(dy.p, y_canonical, static_cast<size_t>(ctx->prob.m) * 8, hipMemcpyHostToDevice));
    return audit_device(ctx, dx.as<double>(), dy.as<double>(), out8);
}

int vj_kernel_stats(vj_ctx* ctx, vj_kernel_stat* out, int cap) {
    if (!ctx || !out || cap < 0) return 0;
    int k = 0;
    for (const auto& t : ctx->timers) {
        if (k >= cap) break;
        std::snprintf(out[k].name, sizeof(out[k].name), "%s", t.name.c_str());
        out[k].avg_us = t.launches ? 1000.0 * t.total_ms / t.launches : 0.0;
        out[k].launches = t.launches;
        out[k].bytes_per_launch = t.bytes;
        ++k;
    }
    return k;
}

int vj_benchmark_kernels(vj_ctx* ctx, int reps, vj_kernel_stat* out, int cap) {
    clear_error();
    if (!ctx || !out || reps <= 0) return -VJ_ERR_INVALID_ARGUMENT;
    if (hipSetDevice(ctx->device) != hipSuccess) return -VJ_ERR_LAUNCH;
    int written = 0;
    const int rc = fast_benchmark(ctx, reps, out, cap, &written);
    if (rc != VJ_OK) return -rc;
    ctx->timers.clear();
    for (int k = 0; k < written; ++k) {
        KernelTimer t;
        t.name = out[k].name;
        t.total_ms = out[k].avg_us * reps / 1000.0;
        t.launches = reps;
        t.bytes = out[k].bytes_per_launch;
        ctx->timers.push_back(t);
    }
    return written;
}

int vj_scaling_factors(vj_ctx* ctx, double* row_scale, double* col_scale) {
    clear_error();
    if (!ctx) return VJ_ERR_INVALID_ARGUMENT;
    VJ_HIP(hipSetDevice(ctx->device));
    return fast_scaling(ctx, row_scale, col_scale);
}

int vj_benchmark_blocks(vj_ctx* ctx, int warmup, int steps, double* ms_per_block, int* iterations_per_block) {
    clear_error();
    if (!ctx || !ms_per_block || !iterations_per_block || warmup < 0 || steps <= 0) return VJ_ERR_INVALID_ARGUMENT;
    if (ctx->params.mode != VJ_MODE_RESTARTED) {
        set_error("check blocks exist in VJ_MODE_RESTARTED");
        return VJ_ERR_UNSUPPORTED;
    }
    VJ_HIP(hipSetDevice(ctx->device));
    return fast_benchmark_blocks(ctx, warmup, steps, ms_per_block, iterations_per_block);
}

}  // extern "C"
