// fv_biot_savart.hip -- line-vortex Biot-Savart sums (collapsed / expanded,
// ring / chordwise / horseshoe legs) and fused AIC assembly.
//
// Reference: warp_kernels.py:25-382 (one thread per (point, vortex) pair, one
// launch per leg, atomic accumulation).  Here the legs of all vortices are
// preprocessed once into a segment table and every target's sum runs in one
// tiled, atomic-free pass.
#include <cmath>
#include <cstdio>
#include <limits>
#include <map>
#include <mutex>

#include "fv_launch.h"

namespace fv {

__global__ void k_fill(double *__restrict__ p, int64_t n, double v)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_fill_i32(int32_t *__restrict__ p, int64_t n, int32_t v)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

int fill(fv_ctx *ctx, double *p, int64_t n, double v)
{
    if (n <= 0) return FV_OK;
    const int b = 256;
    hipLaunchKernelGGL(k_fill, dim3((unsigned)ceil_div(n, b)), dim3(b), 0, ctx->stream, p, n,
                       v);
    FV_LAUNCH_CHECK(ctx);
    return FV_OK;
}

// One thread per (vortex, selected leg).  Corner k of ring m is
// corners[k][3m..3m+2] with corners = (Br, Fr, Fl, Bl); leg k runs k -> k+1.
__global__ void k_build_segtab(const double *__restrict__ br, const double *__restrict__ fr,
                               const double *__restrict__ fl, const double *__restrict__ bl,
                               int64_t m, const double *__restrict__ gamma,
                               const double *__restrict__ rc0, const double *__restrict__ age,
                               double nu, int mask, int nl, int unit_strength,
                               SegRec *__restrict__ out)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= m * nl) return;
    const int64_t v = e / nl;
    const int slot = (int)(e - v * nl);
    int leg = 0, seen = -1;
    for (int k = 0; k < 4; ++k) {
        if (mask & (1 << k)) {
            ++seen;
            if (seen == slot) {
                leg = k;
                break;
            }
        }
    }
    const double *c[4] = {br, fr, fl, bl};
    const double *s = c[leg] + 3 * v;
    const double *t = c[(leg + 1) & 3] + 3 * v;
    SegRec r;
    r.sx = s[0]; r.sy = s[1]; r.sz = s[2];
    r.ex = t[0]; r.ey = t[1]; r.ez = t[2];
    const double r0x = t[0] - s[0], r0y = t[1] - s[1], r0z = t[2] - s[2];
    const double r0sq = r0x * r0x + r0y * r0y + r0z * r0z;
    const double r0len = sqrt(r0sq);
    const double strength = unit_strength ? 1.0 : gamma[v];
    const double a = (unit_strength || !age) ? 0.0 : age[v];
    const double rc = rc0[v];
    // Lamb-Oseen + Squire core growth, warp_kernels.py:88
    const double rcsq = rc * rc + kFourLamb * (nu + kSquire * fabs(strength)) * a;
    r.c1 = strength * kInv4Pi;
    r.c2 = r0sq * rcsq;
    r.tol = (r0len < 2.22e-16) ? INFINITY : r0len * 1.0e-10;
    r.pad = 0.0;
    out[e] = r;
}

int build_segtab(fv_ctx *ctx, const double *br, const double *fr, const double *fl,
                 const double *bl, int64_t m, const double *gamma, const double *rc0,
                 const double *age, double nu, int mask, bool unit_strength, SegRec *out)
{
    const int nl = legs_in_mask(mask);
    const int64_t total = m * nl;
    if (total <= 0) return FV_OK;
    const int b = 256;
    hipLaunchKernelGGL(k_build_segtab, dim3((unsigned)ceil_div(total, b)), dim3(b), 0,
                       ctx->stream, br, fr, fl, bl, m, gamma, rc0, age, nu, mask, nl,
                       unit_strength ? 1 : 0, out);
    FV_LAUNCH_CHECK(ctx);
    return FV_OK;
}

PairPlan plan_pairs(fv_ctx *ctx, int64_t n_tgt, int64_t n_src)
{
    // the shapes the autotuner picks on MI355X for these size classes
    return plan_pairs_shape(ctx, n_tgt, n_src, (n_tgt >= 65536) ? 256 : 64,
                            (n_tgt >= 65536) ? 4 : (n_tgt >= 16384) ? 2 : 1, 16);
}

PairPlan plan_pairs_shape(fv_ctx *ctx, int64_t n_tgt, int64_t n_src, int block, int tpt,
                          int per_cu)
{
    PairPlan p;
    p.block = block;
    p.tpt = tpt;
    const int64_t per_block = (int64_t)p.block * p.tpt;
    p.gx = ceil_div(n_tgt > 0 ? n_tgt : 1, per_block);
    const int64_t want = (int64_t)per_cu * num_cus(ctx);
    int64_t tiles = ceil_div(n_src > 0 ? n_src : 1, p.block);
    int64_t split = ceil_div(want, p.gx);
    if (split > tiles) split = tiles;
    if (split < 1) split = 1;
    if (split > 65535) split = 65535;
    p.chunk = ceil_div(tiles, split) * p.block;
    p.split = (int)ceil_div(n_src > 0 ? n_src : 1, p.chunk);
    return p;
}

template <class Op, int BLOCK, int TPT>
static void launch_pairs(fv_ctx *ctx, const PairPlan &p, const typename Op::Args &a,
                         int64_t n_tgt, int64_t n_src, const int64_t *dyn_tgt,
                         const int64_t *dyn_src, double *dst)
{
    hipLaunchKernelGGL((k_pairs<Op, BLOCK, TPT>), dim3((unsigned)p.gx, (unsigned)p.split),
                       dim3(BLOCK), 0, ctx->stream, a, n_tgt, n_src, dyn_tgt, dyn_src, p.chunk,
                       n_tgt, dst);
}

template <class Op>
static void dispatch_pairs(fv_ctx *ctx, const PairPlan &p, const typename Op::Args &a,
                           int64_t n_tgt, int64_t n_src, const int64_t *dyn_tgt,
                           const int64_t *dyn_src, double *dst)
{
    if (p.block == 256) {
        if (p.tpt == 4) launch_pairs<Op, 256, 4>(ctx, p, a, n_tgt, n_src, dyn_tgt, dyn_src, dst);
        else if (p.tpt == 2) launch_pairs<Op, 256, 2>(ctx, p, a, n_tgt, n_src, dyn_tgt, dyn_src, dst);
        else launch_pairs<Op, 256, 1>(ctx, p, a, n_tgt, n_src, dyn_tgt, dyn_src, dst);
    } else {
        if (p.tpt == 4) launch_pairs<Op, 64, 4>(ctx, p, a, n_tgt, n_src, dyn_tgt, dyn_src, dst);
        else if (p.tpt == 2) launch_pairs<Op, 64, 2>(ctx, p, a, n_tgt, n_src, dyn_tgt, dyn_src, dst);
        else launch_pairs<Op, 64, 1>(ctx, p, a, n_tgt, n_src, dyn_tgt, dyn_src, dst);
    }
}

// ---------------------------------------------------------------- autotuner
// With fv_ctx_set_autotune on, the first pair sum of a kind (operator, device,
// power-of-two size class of targets and sources) times every candidate tile
// shape -- 64 / 256 lanes, 1 / 2 / 4 targets per lane, 4 / 8 / 16 workgroups
// per CU -- with HIP events (each run twice, the second counted), keeps the
// fastest and uses it for every later call of that kind.  The shapes differ
// only in how the sources are split, so results agree to round-off.
struct TuneKey {
    const void *op;
    int device, lt, ls;
    bool operator<(const TuneKey &o) const
    {
        if (op != o.op) return op < o.op;
        if (device != o.device) return device < o.device;
        if (lt != o.lt) return lt < o.lt;
        return ls < o.ls;
    }
};
struct TuneVal {
    int block, tpt, per_cu;
    double ms;
    const char *family;
    int prec;
};
static std::map<TuneKey, TuneVal> g_tune;
static std::mutex g_tune_mu;

static int ilog2_ceil(int64_t n)
{
    int l = 0;
    while ((int64_t(1) << l) < n) ++l;
    return l;
}

template <class Op>
static int tuned_plan(fv_ctx *ctx, const typename Op::Args &a, int64_t n_tgt, int64_t n_src,
                      const int64_t *dyn_tgt, const int64_t *dyn_src, int pslot,
                      const char *family, PairPlan &out)
{
    out = plan_pairs(ctx, n_tgt, n_src);
    if (ctx->fixed_plans && dyn_src) {
        out.chunk = 0;   // sized in the kernel from the device count
        return FV_OK;
    }
    if (!ctx->autotune) return FV_OK;
    // tuning runs 36 trial sums: beyond 2^36 pairs (~50 ms a trial) the
    // default shape (what the tuner picks at 2^17 x 2^17) is kept
    if ((double)n_tgt * (double)n_src > 68719476736.0) return FV_OK;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(ctx->stream, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone)
        return FV_OK;
    const TuneKey key{reinterpret_cast<const void *>(&k_pairs<Op, 64, 1>), ctx->device,
                      ilog2_ceil(n_tgt), ilog2_ceil(n_src > 0 ? n_src : 1)};
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        auto it = g_tune.find(key);
        if (it != g_tune.end()) {
            out = plan_pairs_shape(ctx, n_tgt, n_src, it->second.block, it->second.tpt,
                                   it->second.per_cu);
            return FV_OK;
        }
    }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    FV_HIP_CHECK(ctx, hipEventCreate(&e0));
    FV_HIP_CHECK(ctx, hipEventCreate(&e1));
    TuneVal best{out.block, out.tpt, 16, INFINITY, family, Op::kPrec};
    static const int kBlocks[2] = {64, 256}, kTpt[3] = {1, 2, 4}, kPerCu[3] = {4, 8, 16};
    int rc = FV_OK;
    for (int b : kBlocks)
        for (int t : kTpt)
            for (int w : kPerCu) {
                const PairPlan p = plan_pairs_shape(ctx, n_tgt, n_src, b, t, w);
                double *dst = static_cast<double *>(
                    workspace(ctx, pslot, sizeof(double) * (size_t)p.split * n_tgt * Op::kComp));
                if (!dst) { rc = FV_E_HIP; break; }
                float ms = 0.0f;
                for (int rep = 0; rep < 2 && !rc; ++rep) {
                    if (hipEventRecord(e0, ctx->stream) != hipSuccess) rc = FV_E_HIP;
                    dispatch_pairs<Op>(ctx, p, a, n_tgt, n_src, dyn_tgt, dyn_src, dst);
                    if (hipGetLastError() != hipSuccess) rc = FV_E_HIP;
                    if (hipEventRecord(e1, ctx->stream) != hipSuccess) rc = FV_E_HIP;
                    if (hipEventSynchronize(e1) != hipSuccess) rc = FV_E_HIP;
                    if (!rc && hipEventElapsedTime(&ms, e0, e1) != hipSuccess) rc = FV_E_HIP;
                }
                if (rc) break;
                if (ms < best.ms) best = TuneVal{b, t, w, (double)ms, family, Op::kPrec};
            }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return set_error(ctx, rc, "autotuning the %s pair sum failed", family);
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        g_tune[key] = best;
    }
    out = plan_pairs_shape(ctx, n_tgt, n_src, best.block, best.tpt, best.per_cu);
    return FV_OK;
}

// Source slices from which a wavefront per output adds the partial sums.
constexpr int kWideReduce = 128;

template <class Op>
static int run_pairs(fv_ctx *ctx, const typename Op::Args &a, int64_t n_tgt, int64_t n_src,
                     const int64_t *dyn_tgt, const int64_t *dyn_src, double *out, int pslot,
                     const char *family)
{
    if (n_tgt <= 0) return FV_OK;
    PairPlan p;
    int rc = tuned_plan<Op>(ctx, a, n_tgt, n_src, dyn_tgt, dyn_src, pslot, family, p);
    if (rc) return rc;
    double *dst = out;
    if (p.split > 1) {
        dst = static_cast<double *>(
            workspace(ctx, pslot, sizeof(double) * (size_t)p.split * n_tgt * Op::kComp));
        if (!dst) return FV_E_HIP;
    }
    {
        TimedLaunch tl(ctx, family);
        dispatch_pairs<Op>(ctx, p, a, n_tgt, n_src, dyn_tgt, dyn_src, dst);
        FV_LAUNCH_CHECK(ctx);
        tl.done();
    }
    if (p.split > 1) {
        const int64_t tot = n_tgt * Op::kComp;
        // few targets over many sources (the particle field at the collocation
        // points): a wavefront per output instead of one lane walking the slices
        if (p.split >= kWideReduce)
            hipLaunchKernelGGL((k_reduce_partials_wide<Op::kComp>),
                               dim3((unsigned)ceil_div(tot, 4)), dim3(256), 0, ctx->stream, dst,
                               n_tgt, dyn_tgt, n_tgt, p.split, out);
        else
            hipLaunchKernelGGL((k_reduce_partials<Op::kComp>), dim3((unsigned)ceil_div(tot, 256)),
                               dim3(256), 0, ctx->stream, dst, n_tgt, dyn_tgt, n_tgt, p.split,
                               out);
        FV_LAUNCH_CHECK(ctx);
    }
    return FV_OK;
}

int seg_sum(fv_ctx *ctx, const double *pts, int64_t n, const SegRec *segs, int64_t nseg,
            int precision, double *out, int pslot)
{
    if (n <= 0) return FV_OK;
    if (nseg <= 0) return fill(ctx, out, 3 * n, 0.0);
    if (precision == FV_PREC_FP64) {
        SegOp<0>::Args a{pts, segs};
        return run_pairs<SegOp<0>>(ctx, a, n, nseg, nullptr, nullptr, out, pslot, "seg_pairs");
    }
    SegOp<1>::Args a{pts, segs};
    return run_pairs<SegOp<1>>(ctx, a, n, nseg, nullptr, nullptr, out, pslot, "seg_pairs");
}

template <int KERNEL, int PREC, int MODE>
static int part_sum_t(fv_ctx *ctx, const PartSrcArgs &src, const PartTgtArgs &tgt,
                      int64_t n_tgt, int64_t n_src, const int64_t *dyn_tgt,
                      const int64_t *dyn_src, double *out, int pslot, const char *family)
{
    using Op = PartOp<KERNEL, PREC, MODE>;
    typename Op::Args a{src, tgt};
    return run_pairs<Op>(ctx, a, n_tgt, n_src, dyn_tgt, dyn_src, out, pslot, family);
}

int part_sum(fv_ctx *ctx, int kernel, int precision, int mode, const PartSrcArgs &src,
             const PartTgtArgs &tgt, int64_t n_tgt, int64_t n_src, const int64_t *dyn_tgt,
             const int64_t *dyn_src, double *out, int pslot, const char *family)
{
    static const char *const kFamilies[4] = {"particle_vel", "particle_jac", "particle_stretch",
                                             "particle_curl"};
    if (!family) family = kFamilies[mode == 4 ? 3 : mode & 3];
    if (n_tgt <= 0) return FV_OK;
    const int comp = (mode == 1) ? 9 : (mode == 4) ? 6 : 3;
    if (n_src <= 0) return fill(ctx, out, comp * n_tgt, 0.0);
    PartSrcArgs s = src;
    s.kernel = kernel;
#define FV_PS(K, P, M) \
    return part_sum_t<K, P, M>(ctx, s, tgt, n_tgt, n_src, dyn_tgt, dyn_src, out, pslot, family)
#define FV_PS_MODE(K, P)        \
    switch (mode) {             \
    case 0: FV_PS(K, P, 0);     \
    case 1: FV_PS(K, P, 1);     \
    case 2: FV_PS(K, P, 2);     \
    default: FV_PS(K, P, 3);    \
    }
    if (mode == 4) {   // Gaussian vorticity under the reference CPU pair rule
        if (kernel != FV_KERNEL_GAUSSIAN_ERF)
            return set_error(ctx, FV_E_ARG, "mode 4 is the Gaussian field's vorticity");
        if (precision == FV_PREC_FP64) FV_PS(0, 0, 4);
        FV_PS(0, 1, 4);
    }
    if (kernel == FV_KERNEL_WINCKELMANS) {
        if (precision == FV_PREC_FP64) { FV_PS_MODE(1, 0) }
        else { FV_PS_MODE(1, 1) }
    } else {
        if (precision == FV_PREC_FP64) { FV_PS_MODE(0, 0) }
        else { FV_PS_MODE(0, 1) }
    }
#undef FV_PS_MODE
#undef FV_PS
}

// Expanded: one thread per (point, vortex); legs of the vortex summed in
// table order.  Output (n, m, 3), 24 B per pair written once.
__global__ void k_expanded(const double *__restrict__ pts, int64_t n,
                           const SegRec *__restrict__ segs, int64_t m, int nl,
                           double *__restrict__ out)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n * m) return;
    const int64_t i = e / m, v = e - i * m;
    const double px = pts[3 * i], py = pts[3 * i + 1], pz = pts[3 * i + 2];
    double ax = 0.0, ay = 0.0, az = 0.0;
    for (int k = 0; k < nl; ++k) seg_pair_f64(px, py, pz, segs[v * nl + k], ax, ay, az);
    out[3 * e] = ax;
    out[3 * e + 1] = ay;
    out[3 * e + 2] = az;
}

// AIC: one thread per (collocation point i, ring j); 4 unit-strength legs,
// projected on the unit normal of panel i.
__global__ void k_aic(const double *__restrict__ colloc, const double *__restrict__ normals,
                      int64_t p, const SegRec *__restrict__ segs, double *__restrict__ aic)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= p * p) return;
    const int64_t i = e / p, j = e - i * p;
    const double px = colloc[3 * i], py = colloc[3 * i + 1], pz = colloc[3 * i + 2];
    double ax = 0.0, ay = 0.0, az = 0.0;
#pragma unroll
    for (int k = 0; k < 4; ++k) seg_pair_f64(px, py, pz, segs[4 * j + k], ax, ay, az);
    aic[e] = ax * normals[3 * i] + ay * normals[3 * i + 1] + az * normals[3 * i + 2];
}

int aic_assemble(fv_ctx *ctx, const double *colloc, const double *normals, int64_t p,
                 const SegRec *unit_segs, double *aic)
{
    if (p <= 0) return FV_OK;
    const int b = 256;
    TimedLaunch tl(ctx, "aic");
    hipLaunchKernelGGL(k_aic, dim3((unsigned)ceil_div(p * p, b)), dim3(b), 0, ctx->stream,
                       colloc, normals, p, unit_segs, aic);
    FV_LAUNCH_CHECK(ctx);
    tl.done();
    return FV_OK;
}

}  // namespace fv

using namespace fv;

static int check_vortex_args(fv_ctx *ctx, const double *points, int64_t n, const double *br,
                             const double *fr, const double *fl, const double *bl, int64_t m,
                             const double *strengths, const double *r_c0s, int32_t leg_mask,
                             int32_t precision, const void *out)
{
    if (!ctx) return FV_E_ARG;
    FV_REQUIRE(ctx, n >= 0 && m >= 0, "negative size (n=%lld, m=%lld)", (long long)n,
               (long long)m);
    FV_REQUIRE(ctx, leg_mask > 0 && leg_mask <= 0xF, "leg_mask %d outside 1..15", leg_mask);
    FV_REQUIRE(ctx, precision == FV_PREC_FP64 || precision == FV_PREC_MIXED,
               "unknown precision %d", precision);
    FV_REQUIRE(ctx, n == 0 || (points && out), "null points/out");
    FV_REQUIRE(ctx, m == 0 || (br && fr && fl && bl && strengths && r_c0s),
               "null vortex arrays");
    return FV_OK;
}

// ------------------------------------------------ no-allocation collapsed sum
// fv_vortex_collapsed_dev_noalloc keeps its segment table and per-split partial
// sums in this file-scope device array instead of context workspaces, so it
// never allocates.  Each device the library runs on holds its own copy until
// the process ends.  In host code `scratch` only names the symbol: its device
// address comes from hipGetSymbolAddress, looked up once per device (with that
// device current) at the entry point's first call there, and kept below.
constexpr size_t kNoAllocScratchBytes = size_t(32) << 20;   // 32 MiB per device
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];
constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void *g_scratch_dev[kMaxDevices];

static int noalloc_scratch(fv_ctx *ctx, void **out)
{
    FV_REQUIRE(ctx, ctx->device >= 0 && ctx->device < kMaxDevices,
               "device %d outside the scratch table", ctx->device);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_dev[ctx->device]) {
        int cur = -1;
        FV_HIP_CHECK(ctx, hipGetDevice(&cur));
        if (cur != ctx->device) FV_HIP_CHECK(ctx, hipSetDevice(ctx->device));
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (cur != ctx->device) (void)hipSetDevice(cur);
        if (e != hipSuccess || !ptr)
            return set_error(ctx, FV_E_HIP, "hipGetSymbolAddress(scratch) failed: %s",
                             hipGetErrorString(e));
        g_scratch_dev[ctx->device] = ptr;
    }
    *out = g_scratch_dev[ctx->device];
    return FV_OK;
}

extern "C" {

int fv_vortex_collapsed_dev(fv_ctx *ctx, const double *points, int64_t n, const double *br,
                            const double *fr, const double *fl, const double *bl, int64_t m,
                            const double *strengths, const double *r_c0s, const double *ages,
                            double nu, int32_t leg_mask, int32_t precision, double *out)
{
    int rc = check_vortex_args(ctx, points, n, br, fr, fl, bl, m, strengths, r_c0s, leg_mask,
                               precision, out);
    if (rc) return rc;
    if (n == 0) return FV_OK;
    if (m == 0) return fill(ctx, out, 3 * n, 0.0);
    const int nl = legs_in_mask(leg_mask);
    SegRec *tab = static_cast<SegRec *>(workspace(ctx, S_SEGTAB, sizeof(SegRec) * m * nl));
    if (!tab) return FV_E_HIP;
    rc = build_segtab(ctx, br, fr, fl, bl, m, strengths, r_c0s, ages, nu, leg_mask, false, tab);
    if (rc) return rc;
    return seg_sum(ctx, points, n, tab, m * nl, precision, out, S_PARTIAL);
}

int fv_vortex_collapsed_dev_noalloc(fv_ctx *ctx, const double *points, int64_t n,
                                    const double *br, const double *fr, const double *fl,
                                    const double *bl, int64_t m, const double *strengths,
                                    const double *r_c0s, const double *ages, double nu,
                                    int32_t leg_mask, int32_t precision, double *out)
{
    int rc = check_vortex_args(ctx, points, n, br, fr, fl, bl, m, strengths, r_c0s, leg_mask,
                               precision, out);
    if (rc) return rc;
    if (n == 0) return FV_OK;
    if (m == 0) return fill(ctx, out, 3 * n, 0.0);
    const int nl = legs_in_mask(leg_mask);
    const int64_t nseg = m * nl;
    const size_t seg_bytes = (sizeof(SegRec) * (size_t)nseg + 255) & ~size_t(255);
    FV_REQUIRE(ctx, seg_bytes <= kNoAllocScratchBytes,
               "no-allocation sum: %lld segments need %zu bytes of segment table, over the "
               "%zu-byte scratch (at most %zu segments)", (long long)nseg, seg_bytes,
               kNoAllocScratchBytes, kNoAllocScratchBytes / sizeof(SegRec));
    void *base = nullptr;
    rc = noalloc_scratch(ctx, &base);
    if (rc) return rc;
    SegRec *tab = static_cast<SegRec *>(base);
    double *part = reinterpret_cast<double *>(static_cast<unsigned char *>(base) + seg_bytes);
    // the default tile shape (that of fv_vortex_collapsed_dev without the
    // autotuner), its source split capped so the partial sums fit
    PairPlan p = plan_pairs(ctx, n, nseg);
    const int64_t max_split =
        (int64_t)((kNoAllocScratchBytes - seg_bytes) / (sizeof(double) * 3 * (size_t)n));
    if (p.split > 1 && p.split > max_split) {
        const int64_t tiles = ceil_div(nseg, p.block);
        const int64_t cap = max_split > 1 ? max_split : 1;
        p.chunk = ceil_div(tiles, cap) * p.block;
        p.split = (int)ceil_div(nseg, p.chunk);
    }
    rc = build_segtab(ctx, br, fr, fl, bl, m, strengths, r_c0s, ages, nu, leg_mask, false, tab);
    if (rc) return rc;
    double *dst = p.split > 1 ? part : out;
    if (precision == FV_PREC_FP64) {
        SegOp<0>::Args a{points, tab};
        dispatch_pairs<SegOp<0>>(ctx, p, a, n, nseg, nullptr, nullptr, dst);
    } else {
        SegOp<1>::Args a{points, tab};
        dispatch_pairs<SegOp<1>>(ctx, p, a, n, nseg, nullptr, nullptr, dst);
    }
    FV_LAUNCH_CHECK(ctx);
    if (p.split > 1) {
        // the reduction run_pairs picks for this split
        if (p.split >= kWideReduce)
            hipLaunchKernelGGL((k_reduce_partials_wide<3>), dim3((unsigned)ceil_div(3 * n, 4)),
                               dim3(256), 0, ctx->stream, part, n, (const int64_t *)nullptr, n,
                               p.split, out);
        else
            hipLaunchKernelGGL((k_reduce_partials<3>), dim3((unsigned)ceil_div(3 * n, 256)),
                               dim3(256), 0, ctx->stream, part, n, (const int64_t *)nullptr, n,
                               p.split, out);
        FV_LAUNCH_CHECK(ctx);
    }
    return FV_OK;
}

int fv_vortex_collapsed_host(fv_ctx *ctx, const double *points, int64_t n, const double *br,
                             const double *fr, const double *fl, const double *bl, int64_t m,
                             const double *strengths, const double *r_c0s, const double *ages,
                             double nu, int32_t leg_mask, int32_t precision, double *out)
{
    int rc = check_vortex_args(ctx, points, n, br, fr, fl, bl, m, strengths, r_c0s, leg_mask,
                               precision, out);
    if (rc) return rc;
    if (n == 0) return FV_OK;
    if (m == 0) {
        for (int64_t i = 0; i < 3 * n; ++i) out[i] = 0.0;
        return FV_OK;
    }
    const double *dp = upload(ctx, S_POINTS, points, 3 * n);
    const double *dbr = upload(ctx, S_BR, br, 3 * m);
    const double *dfr = upload(ctx, S_FR, fr, 3 * m);
    const double *dfl = upload(ctx, S_FL, fl, 3 * m);
    const double *dbl = upload(ctx, S_BL, bl, 3 * m);
    const double *dg = upload(ctx, S_GAMMA, strengths, m);
    const double *drc = upload(ctx, S_RC0, r_c0s, m);
    const double *dage = ages ? upload(ctx, S_AGE, ages, m) : nullptr;
    double *dout = static_cast<double *>(workspace(ctx, S_OUT, sizeof(double) * 3 * n));
    if (!dp || !dbr || !dfr || !dfl || !dbl || !dg || !drc || (ages && !dage) || !dout)
        return FV_E_HIP;
    rc = fv_vortex_collapsed_dev(ctx, dp, n, dbr, dfr, dfl, dbl, m, dg, drc, dage, nu, leg_mask,
                                 precision, dout);
    if (rc) return rc;
    return fv_memcpy_d2h(ctx, out, dout, sizeof(double) * 3 * n);
}

int fv_vortex_expanded_dev(fv_ctx *ctx, const double *points, int64_t n, const double *br,
                           const double *fr, const double *fl, const double *bl, int64_t m,
                           const double *strengths, const double *r_c0s, const double *ages,
                           double nu, int32_t leg_mask, double *out)
{
    int rc = check_vortex_args(ctx, points, n, br, fr, fl, bl, m, strengths, r_c0s, leg_mask,
                               FV_PREC_FP64, out);
    if (rc) return rc;
    if (n == 0 || m == 0) return FV_OK;
    const int nl = legs_in_mask(leg_mask);
    SegRec *tab = static_cast<SegRec *>(workspace(ctx, S_SEGTAB, sizeof(SegRec) * m * nl));
    if (!tab) return FV_E_HIP;
    rc = build_segtab(ctx, br, fr, fl, bl, m, strengths, r_c0s, ages, nu, leg_mask, false, tab);
    if (rc) return rc;
    const int b = 256;
    TimedLaunch tl(ctx, "seg_expanded");
    hipLaunchKernelGGL(k_expanded, dim3((unsigned)ceil_div(n * m, b)), dim3(b), 0, ctx->stream,
                       points, n, tab, m, nl, out);
    FV_LAUNCH_CHECK(ctx);
    tl.done();
    return FV_OK;
}

int fv_vortex_expanded_host(fv_ctx *ctx, const double *points, int64_t n, const double *br,
                            const double *fr, const double *fl, const double *bl, int64_t m,
                            const double *strengths, const double *r_c0s, const double *ages,
                            double nu, int32_t leg_mask, double *out)
{
    int rc = check_vortex_args(ctx, points, n, br, fr, fl, bl, m, strengths, r_c0s, leg_mask,
                               FV_PREC_FP64, out);
    if (rc) return rc;
    if (n == 0 || m == 0) return FV_OK;
    const double *dp = upload(ctx, S_POINTS, points, 3 * n);
    const double *dbr = upload(ctx, S_BR, br, 3 * m);
    const double *dfr = upload(ctx, S_FR, fr, 3 * m);
    const double *dfl = upload(ctx, S_FL, fl, 3 * m);
    const double *dbl = upload(ctx, S_BL, bl, 3 * m);
    const double *dg = upload(ctx, S_GAMMA, strengths, m);
    const double *drc = upload(ctx, S_RC0, r_c0s, m);
    const double *dage = ages ? upload(ctx, S_AGE, ages, m) : nullptr;
    double *dout = static_cast<double *>(workspace(ctx, S_OUT, sizeof(double) * 3 * n * m));
    if (!dp || !dbr || !dfr || !dfl || !dbl || !dg || !drc || (ages && !dage) || !dout)
        return FV_E_HIP;
    rc = fv_vortex_expanded_dev(ctx, dp, n, dbr, dfr, dfl, dbl, m, dg, drc, dage, nu, leg_mask,
                                dout);
    if (rc) return rc;
    return fv_memcpy_d2h(ctx, out, dout, sizeof(double) * 3 * n * m);
}

int fv_aic_dev(fv_ctx *ctx, const double *colloc, const double *normals, int64_t p,
               const double *br, const double *fr, const double *fl, const double *bl,
               const double *r_c0s, double nu, double *aic)
{
    if (!ctx) return FV_E_ARG;
    FV_REQUIRE(ctx, p >= 0, "negative panel count");
    if (p == 0) return FV_OK;
    FV_REQUIRE(ctx, colloc && normals && br && fr && fl && bl && r_c0s && aic,
               "null AIC argument");
    SegRec *tab = static_cast<SegRec *>(workspace(ctx, S_SEGTAB, sizeof(SegRec) * p * 4));
    if (!tab) return FV_E_HIP;
    int rc = build_segtab(ctx, br, fr, fl, bl, p, nullptr, r_c0s, nullptr, nu, FV_LEGS_RING,
                          true, tab);
    if (rc) return rc;
    return aic_assemble(ctx, colloc, normals, p, tab, aic);
}

int fv_aic_host(fv_ctx *ctx, const double *colloc, const double *normals, int64_t p,
                const double *br, const double *fr, const double *fl, const double *bl,
                const double *r_c0s, double nu, double *aic)
{
    if (!ctx) return FV_E_ARG;
    FV_REQUIRE(ctx, p >= 0, "negative panel count");
    if (p == 0) return FV_OK;
    FV_REQUIRE(ctx, colloc && normals && br && fr && fl && bl && r_c0s && aic,
               "null AIC argument");
    const double *dc = upload(ctx, S_POINTS, colloc, 3 * p);
    const double *dn = upload(ctx, S_NORMALS, normals, 3 * p);
    const double *dbr = upload(ctx, S_BR, br, 3 * p);
    const double *dfr = upload(ctx, S_FR, fr, 3 * p);
    const double *dfl = upload(ctx, S_FL, fl, 3 * p);
    const double *dbl = upload(ctx, S_BL, bl, 3 * p);
    const double *drc = upload(ctx, S_RC0, r_c0s, p);
    double *dout = static_cast<double *>(workspace(ctx, S_OUT, sizeof(double) * p * p));
    if (!dc || !dn || !dbr || !dfr || !dfl || !dbl || !drc || !dout) return FV_E_HIP;
    int rc = fv_aic_dev(ctx, dc, dn, p, dbr, dfr, dfl, dbl, drc, nu, dout);
    if (rc) return rc;
    return fv_memcpy_d2h(ctx, aic, dout, sizeof(double) * p * p);
}

}  // extern "C"

extern "C" {

int fv_ctx_set_autotune(fv_ctx *ctx, int32_t enabled)
{
    if (!ctx) return FV_E_ARG;
    ctx->autotune = enabled != 0;
    return FV_OK;
}

// One line per tuned kind: family (+ "_mixed" for FP32 pair terms), size
// classes (log2 targets, sources),
// lanes per workgroup, targets per lane, workgroups per CU, best time (ms).
int fv_autotune_report(fv_ctx *ctx, char *buf, int64_t cap)
{
    if (!ctx || !buf || cap <= 0) return FV_E_ARG;
    std::lock_guard<std::mutex> lk(fv::g_tune_mu);
    int64_t used = 0;
    buf[0] = 0;
    for (const auto &kv : fv::g_tune) {
        if (kv.first.device != ctx->device) continue;
        const int w = snprintf(buf + used, (size_t)(cap - used), "%s%s %d %d %d %d %d %.4f\n",
                               kv.second.family ? kv.second.family : "?",
                               kv.second.prec ? "_mixed" : "", kv.first.lt, kv.first.ls,
                               kv.second.block, kv.second.tpt, kv.second.per_cu, kv.second.ms);
        if (w < 0 || used + w >= cap) break;
        used += w;
    }
    return FV_OK;
}

}  // extern "C"
