// Internal declarations shared by the HIP translation units of libvajra_hip.so.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "vajra_pdhg.h"

namespace vj {

// ---- error plumbing ---------------------------------------------------------------
void set_error(const std::string& msg);
void clear_error();

#define VJ_HIP(call)                                                                    \
    do {                                                                                \
        hipError_t e_ = (call);                                                         \
        if (e_ != hipSuccess) {                                                         \
            ::vj::set_error(std::string(#call) + " failed: " + hipGetErrorString(e_) +  \
                            " (" + __FILE__ + ":" + std::to_string(__LINE__) + ")");    \
            return VJ_ERR_LAUNCH;                                                       \
        }                                                                               \
    } while (0)

#define VJ_LAUNCHED()                                                                   \
    do {                                                                                \
        hipError_t e_ = hipGetLastError();                                              \
        if (e_ != hipSuccess) {                                                         \
            ::vj::set_error(std::string("kernel launch failed: ") + hipGetErrorString(e_) + \
                            " (" + __FILE__ + ":" + std::to_string(__LINE__) + ")");    \
            return VJ_ERR_LAUNCH;                                                       \
        }                                                                               \
    } while (0)

#define VJ_TRY(...)                         \
    do {                                    \
        int rc_ = (__VA_ARGS__);            \
        if (rc_ != VJ_OK) return rc_;       \
    } while (0)

// ---- device memory owned by a context ---------------------------------------------
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    int alloc(size_t b) {
        release();
        if (b == 0) b = 16;
        hipError_t e = hipMalloc(&p, b);
        if (e != hipSuccess) {
            p = nullptr;
            set_error("hipMalloc(" + std::to_string(b) + " bytes) failed: " + hipGetErrorString(e));
            return VJ_ERR_ALLOC;
        }
        bytes = b;
        return VJ_OK;
    }
    template <class T> T* as() const { return static_cast<T*>(p); }
};

inline int grid_for(int64_t work, int per_block) {
    int64_t g = (work + per_block - 1) / per_block;
    return static_cast<int>(g < 1 ? 1 : g);
}

// ---- per-instance solver state (device memory, one entry per LP) --------------------
// Scalars read by the iteration kernels and written only by the control kernel.
struct alignas(16) InstState {
    double tau, sigma, omega, eta;     // steps: tau = eta/omega, sigma = eta*omega
    double r0, r_prev;                 // fixed-point residual at epoch start / last check
    double rel_kkt, rel_p, rel_d, rel_gap, pobj, dobj;
    double norm_b, norm_c;             // original-space norms for the rel-KKT denominators
    double obj_scale, bound_scale;     // cs, bs of the rescaled problem
    double tol;
    double kkt_min;                    // least rel-KKT at the checks of the current epoch (0: none yet)
    double time_to_target;             // filled by host
    double pw_sum;                     // sum of the primal-weight errors at the restarts so far
    int64_t k;                         // iterations since the last restart (epoch index)
    int64_t iters;                     // total iterations
    int64_t max_iters;
    int32_t active;                    // 1 while iterating
    int32_t restart_now;               // set by control, consumed by restart kernels
    int32_t restarted_prev;            // a restart happened at the end of the previous block
    int32_t restarts;
    int32_t converged;
    int32_t nan_flag;
    int32_t checks;
    int32_t anchored;                  // the centre moved at a check of this epoch: the steps add the anchor term
};

}  // namespace vj
