// fv_runtime.hip -- context, workspaces, errors, copies and timing of
// libfluxvortex_hip.so.
#include <cstring>
#include <new>

#include "fv_common.h"

namespace fv {

int set_error(fv_ctx *ctx, int code, const char *fmt, ...)
{
    if (ctx) {
        char buf[1024];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof buf, fmt, ap);
        va_end(ap);
        ctx->last_error = buf;
    }
    return code;
}

void *workspace(fv_ctx *ctx, int slot, size_t bytes)
{
    if (slot < 0 || slot >= kWorkspaceSlots) {
        set_error(ctx, FV_E_ARG, "workspace slot %d out of range", slot);
        return nullptr;
    }
    Workspace &w = ctx->ws[slot];
    if (bytes == 0) bytes = 16;
    if (w.bytes >= bytes) return w.ptr;
    if (w.ptr) {
        // the stream may still read the old buffer
        (void)hipStreamSynchronize(ctx->stream);
        (void)hipFree(w.ptr);
        w.ptr = nullptr;
        w.bytes = 0;
    }
    size_t cap = bytes + bytes / 4;
    cap = (cap + 255) & ~size_t(255);
    hipError_t e = hipMalloc(&w.ptr, cap);
    if (e != hipSuccess) {
        w.ptr = nullptr;
        set_error(ctx, FV_E_HIP, "hipMalloc(%zu) for workspace %d failed: %s", cap, slot,
                  hipGetErrorString(e));
        return nullptr;
    }
    w.bytes = cap;
    return w.ptr;
}

double *upload(fv_ctx *ctx, int slot, const double *host, size_t count)
{
    double *d = static_cast<double *>(workspace(ctx, slot, count * sizeof(double)));
    if (!d) return nullptr;
    if (count) {
        hipError_t e = hipMemcpyAsync(d, host, count * sizeof(double), hipMemcpyHostToDevice,
                                      ctx->stream);
        if (e != hipSuccess) {
            set_error(ctx, FV_E_HIP, "hipMemcpyAsync H2D failed: %s", hipGetErrorString(e));
            return nullptr;
        }
    }
    return d;
}

static hipEvent_t pooled_event(fv_ctx *ctx)
{
    if (!ctx->event_pool.empty()) {
        hipEvent_t e = ctx->event_pool.back();
        ctx->event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

TimedLaunch::TimedLaunch(fv_ctx *c, const char *fam, hipStream_t s)
    : ctx(c), family(fam), stream(s ? s : c->stream)
{
    if (ctx->timing) {
        start = pooled_event(ctx);
        if (start) (void)hipEventRecord(start, stream);
    }
}

void TimedLaunch::done()
{
    if (!ctx->timing || !start) return;
    hipEvent_t stop = pooled_event(ctx);
    if (!stop) return;
    (void)hipEventRecord(stop, stream);
    ctx->timed.push_back(fv_timed_launch{family, start, stop});
    start = nullptr;
}

int num_cus(fv_ctx *ctx)
{
    if (ctx->cus > 0) return ctx->cus;
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, ctx->device) !=
            hipSuccess ||
        v <= 0)
        v = 256;
    ctx->cus = v;
    return v;
}

}  // namespace fv

using namespace fv;

extern "C" {

int fv_ctx_create(int device, fv_ctx **out)
{
    if (!out) return FV_E_ARG;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return FV_E_HIP;
    if (device < 0 || device >= count) return FV_E_ARG;
    if (hipSetDevice(device) != hipSuccess) return FV_E_HIP;
    fv_ctx *ctx = new (std::nothrow) fv_ctx();
    if (!ctx) return FV_E_HIP;
    ctx->device = device;
    if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
        delete ctx;
        return FV_E_HIP;
    }
    for (auto &e : ctx->user_events) {
        if (hipEventCreate(&e) != hipSuccess) {
            fv_ctx_destroy(ctx);
            return FV_E_HIP;
        }
    }
    num_cus(ctx);
    *out = ctx;
    return FV_OK;
}

void fv_ctx_destroy(fv_ctx *ctx)
{
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    for (auto &w : ctx->ws)
        if (w.ptr) (void)hipFree(w.ptr);
    for (auto &t : ctx->timed) {
        (void)hipEventDestroy(t.start);
        (void)hipEventDestroy(t.stop);
    }
    for (auto e : ctx->event_pool) (void)hipEventDestroy(e);
    for (auto e : ctx->user_events)
        if (e) (void)hipEventDestroy(e);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

const char *fv_last_error(const fv_ctx *ctx)
{
    if (!ctx) return "null context";
    return ctx->last_error.c_str();
}

int fv_ctx_synchronize(fv_ctx *ctx)
{
    if (!ctx) return FV_E_ARG;
    FV_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return FV_OK;
}

void *fv_ctx_stream(fv_ctx *ctx) { return ctx ? (void *)ctx->stream : nullptr; }

int fv_device_alloc(fv_ctx *ctx, size_t bytes, void **out)
{
    if (!ctx || !out) return FV_E_ARG;
    *out = nullptr;
    FV_HIP_CHECK(ctx, hipMalloc(out, bytes ? bytes : 16));
    return FV_OK;
}

int fv_device_free(fv_ctx *ctx, void *ptr)
{
    if (!ctx) return FV_E_ARG;
    if (!ptr) return FV_OK;
    FV_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    FV_HIP_CHECK(ctx, hipFree(ptr));
    return FV_OK;
}

int fv_memcpy_h2d(fv_ctx *ctx, void *dst, const void *src, size_t bytes)
{
    if (!ctx || (bytes && (!dst || !src))) return FV_E_ARG;
    if (!bytes) return FV_OK;
    FV_HIP_CHECK(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    FV_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return FV_OK;
}

int fv_memcpy_d2h(fv_ctx *ctx, void *dst, const void *src, size_t bytes)
{
    if (!ctx || (bytes && (!dst || !src))) return FV_E_ARG;
    if (!bytes) return FV_OK;
    FV_HIP_CHECK(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    FV_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return FV_OK;
}

int fv_ctx_set_timing(fv_ctx *ctx, int enabled)
{
    if (!ctx) return FV_E_ARG;
    FV_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    for (auto &t : ctx->timed) {
        ctx->event_pool.push_back(t.start);
        ctx->event_pool.push_back(t.stop);
    }
    ctx->timed.clear();
    ctx->timing = enabled != 0;
    return FV_OK;
}

int fv_ctx_kernel_stats(fv_ctx *ctx, const char *family, int64_t *launches, double *total_ms)
{
    if (!ctx || !family) return FV_E_ARG;
    FV_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    int64_t n = 0;
    double tot = 0.0;
    for (auto &t : ctx->timed) {
        if (t.family != family) continue;
        float ms = 0.f;
        FV_HIP_CHECK(ctx, hipEventElapsedTime(&ms, t.start, t.stop));
        tot += ms;
        ++n;
    }
    if (launches) *launches = n;
    if (total_ms) *total_ms = tot;
    return FV_OK;
}

int fv_ctx_event_record(fv_ctx *ctx, int32_t which)
{
    if (!ctx || which < 0 || which >= 8) return FV_E_ARG;
    FV_HIP_CHECK(ctx, hipEventRecord(ctx->user_events[which], ctx->stream));
    return FV_OK;
}

int fv_ctx_event_elapsed(fv_ctx *ctx, int32_t a, int32_t b, double *ms)
{
    if (!ctx || !ms || a < 0 || a >= 8 || b < 0 || b >= 8) return FV_E_ARG;
    FV_HIP_CHECK(ctx, hipEventSynchronize(ctx->user_events[b]));
    float f = 0.f;
    FV_HIP_CHECK(ctx, hipEventElapsedTime(&f, ctx->user_events[a], ctx->user_events[b]));
    *ms = f;
    return FV_OK;
}

}  // extern "C"
