// fv_common.h -- shared host-side definitions of libfluxvortex_hip.so:
// the context (stream, grow-only workspaces, last error), status handling and
// launch helpers.  Device math lives in fv_pairs.h.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/fluxvortex.h"

namespace fv {

constexpr int kWorkspaceSlots = 24;

struct Workspace {
    void *ptr = nullptr;
    size_t bytes = 0;
};

}  // namespace fv

struct fv_timed_launch {
    std::string family;
    hipEvent_t start, stop;
};

struct fv_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string last_error;
    fv::Workspace ws[fv::kWorkspaceSlots];
    bool timing = false;
    std::vector<fv_timed_launch> timed;   // launches recorded while timing
    std::vector<hipEvent_t> event_pool;   // recycled events
    hipEvent_t user_events[8] = {};
    int cus = 0;
    bool autotune = false;                // pair-sum tile shapes timed at first use
    // particle pair sums with device counts: the default shape for the upper
    // bound and each source slice's length from the device count (launches
    // that depend on no host-side count, for graph capture)
    bool fixed_plans = false;
};

namespace fv {

// Workspace slot ids (each entry point family uses its own slots so that a
// call never aliases buffers of another in flight on the same stream).
enum Slot : int {
    S_POINTS = 0,
    S_BR, S_FR, S_FL, S_BL,
    S_GAMMA, S_RC0, S_AGE,
    S_SEGTAB,
    S_OUT,
    S_PARTIAL,
    S_NORMALS,
    S_PTGT, S_PPOS, S_PALPHA, S_PSIG,
    S_INFO,
    S_MAT32,
    S_VEC,
    S_TMP0, S_TMP1,
};

int set_error(fv_ctx *ctx, int code, const char *fmt, ...);

// Grow-only workspace; returns nullptr and sets the error on failure.
void *workspace(fv_ctx *ctx, int slot, size_t bytes);

#define FV_HIP_CHECK(ctx, expr)                                                         \
    do {                                                                                \
        hipError_t e_ = (expr);                                                         \
        if (e_ != hipSuccess)                                                           \
            return ::fv::set_error((ctx), FV_E_HIP, "%s failed: %s (%s:%d)", #expr,     \
                                   hipGetErrorString(e_), __FILE__, __LINE__);          \
    } while (0)

#define FV_LAUNCH_CHECK(ctx)                                                            \
    do {                                                                                \
        hipError_t e_ = hipGetLastError();                                              \
        if (e_ != hipSuccess)                                                           \
            return ::fv::set_error((ctx), FV_E_HIP, "kernel launch failed: %s (%s:%d)", \
                                   hipGetErrorString(e_), __FILE__, __LINE__);          \
    } while (0)

#define FV_REQUIRE(ctx, cond, ...)                                                      \
    do {                                                                                \
        if (!(cond)) return ::fv::set_error((ctx), FV_E_ARG, __VA_ARGS__);              \
    } while (0)

// Copies host -> workspace slot; returns device pointer or nullptr.
double *upload(fv_ctx *ctx, int slot, const double *host, size_t count);

// Brackets one launch of a timed kernel family with events when timing is
// on:  TimedLaunch t(ctx, "seg_pairs"); launch...; t.done();  the events go
// on `stream` (default: the context's stream), where the launches go.
struct TimedLaunch {
    fv_ctx *ctx;
    const char *family;
    hipStream_t stream;
    hipEvent_t start = nullptr;
    TimedLaunch(fv_ctx *c, const char *fam, hipStream_t s = nullptr);
    void done();
};

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

int num_cus(fv_ctx *ctx);

}  // namespace fv
