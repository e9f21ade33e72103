// nbh_ctx.h -- the context object behind the C ABI and the helpers shared by
// the translation units that implement it (nbh_api.hip, nbh_ensemble.hip).
#pragma once

#include "nbh_internal.h"
#include "../../include/nbody_hip.h"

#include <cstdint>
#include <string>
#include <vector>

namespace nbh {

// Records msg (printf-style) as this thread's last error and returns code.
int fail(int code, const char* fmt, ...);
const char* last_error();

#define NBH_HIP(call)                                                              \
    do {                                                                           \
        hipError_t e_ = (call);                                                    \
        if (e_ != hipSuccess)                                                      \
            return ::nbh::fail(NBH_ERR_HIP, "%s failed: %s", #call, hipGetErrorString(e_)); \
    } while (0)

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    // Grows to at least `need` bytes (contents not preserved).
    hipError_t reserve(size_t need)
    {
        if (need <= bytes) return hipSuccess;
        release();
        hipError_t e = hipMalloc(&p, need);
        if (e == hipSuccess) bytes = need;
        return e;
    }
    double* d() const { return static_cast<double*>(p); }
};

struct EventPair {
    hipEvent_t a = nullptr, b = nullptr;
};

State state_of(double* base, int64_t n);
Forces forces_of(double* base, int64_t n);

}  // namespace nbh

struct nbh_ctx {
    nbh_config cfg{};
    int cus = 0;
    hipStream_t stream = nullptr;

    // transient buffers of the host-array calls
    nbh::DevBuf in;       // 7n
    nbh::DevBuf out;      // 6n
    nbh::DevBuf partial;  // 6 * jsplit * n
    nbh::DevBuf rec;      // packed source records: 64 B (FP64) / 48 B (mixed) per body
    nbh::DevBuf vwork;    // mixed: reference-velocity reduction, vref_work_doubles()
    void* scratch = nullptr;  // this device's copy of the static scratch (nbh_force_device_noalloc)
    nbh::DevBuf flag;     // int
    nbh::DevBuf scal;     // scalars: [0] dt, [1] dt_next bits, [2..3] energy, [4..] spare

    // resident state
    int n = 0;
    nbh::DevBuf state;  // 7n: x y z m vx vy vz
    nbh::DevBuf f0;     // 6n: forces at the current state
    nbh::DevBuf f1;     // 6n: forces at the predicted state
    nbh::DevBuf ework;  // energy block sums
    bool adaptive = false;
    double eta = 1e-2;
    bool dt_set = false;

    // graph of the step sequence
    hipGraphExec_t gexec = nullptr;
    int g_steps = 0;
    int g_n = -1;
    bool g_adaptive = false;

    // autotuned launch shapes, [precision][floor(log2 n)]: block, ipl, jsplit, source path
    struct Tuned {
        int block = 0, ipl = 0, jsplit = 0, scalar = 0;
    } tuned[2][32];

    // profiling
    bool profiling = false;
    std::vector<nbh::EventPair> ev_free, ev_pending;
    double force_ms = 0.0;
    int64_t force_launches = 0;
    double step_ms = 0.0;
    int64_t steps = 0;
};

namespace nbh {
int get_events(nbh_ctx* c, EventPair* ep);
int settle_events(nbh_ctx* c);

// An event pair taken from the context: handed to the pending list once
// both events are recorded (pending()), else returned to the free list when
// the guard goes out of scope, so that error returns do not lose it.
struct EventGuard {
    nbh_ctx* c = nullptr;
    EventPair ep;
    int take(nbh_ctx* ctx)
    {
        int rc = get_events(ctx, &ep);
        if (rc == 0) c = ctx;
        return rc;
    }
    void pending()
    {
        c->ev_pending.push_back(ep);
        c = nullptr;
    }
    ~EventGuard()
    {
        if (c) c->ev_free.push_back(ep);
    }
};
}  // namespace nbh
