// fv_population.hip -- particle population control and the vector-circulation
// ledger on the device.
//
// Reference: newton_pc/adapters/flap.py
//   drop    :469-473  keep the newest max_particles (the oldest sit at the
//                     front of the arrays);
//   merge   :478-563  pairwise moment-conserving merge (UAV-VPM scheme): pairs
//                     of aligned particles in the same distance-graded hash
//                     cell, accepted when the velocity change at the wing
//                     reference points stays below eps |V_inf|;
//   ledger  :565-578  sum of particle alpha + sum over ring legs of G (b - a).
//
// Merge pipeline (no host synchronisation; sizes from the field's count
// bound, the exact count read on the device):
//   k_pop_keys     candidate flags + 63-bit keys (three 21-bit biased cell
//                  indices, so the integer order is the lexicographic one);
//                  non-candidates get the sentinel ~0
//   radix sort     (key, particle index) pairs, stable (hipCUB): ties keep
//                  the particle order, as a stable lexsort does
//   k_pop_pairs    one thread per run of equal keys walks it greedily
//   k_pop_eval     per formed pair: merged particle, error at the reference
//                  points, accept -> write the merged particle into the
//                  first slot, clear the second's keep flag
//   scan + k_pop_compact   stable compaction of the survivors
// Arithmetic: every expression that decides a key, a pair or an acceptance
// is evaluated without contraction in the order of the CPU statement
// (oracle/fv_oracle_pop.c), so both pick the same pairs.
#include <hipcub/hipcub.hpp>

#include <cmath>
#include <cstring>
#include <new>

#include "fv_vpm.h"

namespace fv {

constexpr int kPopBlock = 256;
constexpr int64_t kKeyBias = int64_t(1) << 20;   // cell indices in (-2^20, 2^20)
constexpr uint64_t kSentinel = ~uint64_t(0);
constexpr int kMaxRef = 64;

struct PopScratch {
    size_t bytes = 0;
    int64_t cap = 0;
    uint64_t *key_in = nullptr, *key_out = nullptr;
    int64_t *val_in = nullptr, *val_out = nullptr;
    int32_t *pair_at = nullptr, *keep = nullptr, *keep_pos = nullptr;
    double *tmp = nullptr;          // 8 cap: compaction staging
    double *ref = nullptr;          // kMaxRef x 3 reference points
    int64_t *counters = nullptr;    // [0] candidates, [1] merged (cumulative), [2] keys out of range
    void *cub = nullptr;
    size_t cub_bytes = 0;
};

static PopScratch *pop_of(fv_vpm *f) { return static_cast<PopScratch *>(f->pop); }

static int pop_ensure(fv_vpm *f)
{
    if (f->pop) return FV_OK;
    fv_ctx *ctx = f->ctx;
    const int64_t cap = f->cap;
    size_t sort_bytes = 0, scan_bytes = 0;
    FV_HIP_CHECK(ctx, hipcub::DeviceRadixSort::SortPairs(nullptr, sort_bytes, (uint64_t *)nullptr,
                                                         (uint64_t *)nullptr, (int64_t *)nullptr,
                                                         (int64_t *)nullptr, (int)cap, 0, 64,
                                                         ctx->stream));
    FV_HIP_CHECK(ctx, hipcub::DeviceScan::ExclusiveSum(nullptr, scan_bytes, (int32_t *)nullptr,
                                                       (int32_t *)nullptr, (int)cap, ctx->stream));
    PopScratch *s = new (std::nothrow) PopScratch();
    if (!s) return set_error(ctx, FV_E_HIP, "out of host memory");
    s->cap = cap;
    s->cub_bytes = sort_bytes > scan_bytes ? sort_bytes : scan_bytes;
    auto al = [](size_t b) { return (b + 255) & ~size_t(255); };
    const size_t n = (size_t)cap;
    size_t off = 0;
    const size_t o_kin = off; off += al(8 * n);
    const size_t o_kout = off; off += al(8 * n);
    const size_t o_vin = off; off += al(8 * n);
    const size_t o_vout = off; off += al(8 * n);
    const size_t o_pair = off; off += al(4 * n);
    const size_t o_keep = off; off += al(4 * n);
    const size_t o_kpos = off; off += al(4 * n);
    const size_t o_tmp = off; off += al(8 * 8 * n);
    const size_t o_ref = off; off += al(8 * 3 * kMaxRef);
    const size_t o_cnt = off; off += al(8 * 4);
    const size_t o_cub = off; off += al(s->cub_bytes);
    void *base = nullptr;
    if (hipMalloc(&base, off) != hipSuccess) {
        delete s;
        return set_error(ctx, FV_E_HIP, "hipMalloc(%zu) for the population scratch failed", off);
    }
    char *b = static_cast<char *>(base);
    s->bytes = off;
    s->key_in = reinterpret_cast<uint64_t *>(b + o_kin);
    s->key_out = reinterpret_cast<uint64_t *>(b + o_kout);
    s->val_in = reinterpret_cast<int64_t *>(b + o_vin);
    s->val_out = reinterpret_cast<int64_t *>(b + o_vout);
    s->pair_at = reinterpret_cast<int32_t *>(b + o_pair);
    s->keep = reinterpret_cast<int32_t *>(b + o_keep);
    s->keep_pos = reinterpret_cast<int32_t *>(b + o_kpos);
    s->tmp = reinterpret_cast<double *>(b + o_tmp);
    s->ref = reinterpret_cast<double *>(b + o_ref);
    s->counters = reinterpret_cast<int64_t *>(b + o_cnt);
    s->cub = b + o_cub;
    f->pop = s;
    const int64_t zero[4] = {0, 0, 0, 0};
    FV_HIP_CHECK(ctx, hipMemcpyAsync(s->counters, zero, sizeof zero, hipMemcpyHostToDevice,
                                     ctx->stream));
    return FV_OK;
}

void pop_destroy(fv_vpm *f)
{
    PopScratch *s = pop_of(f);
    if (!s) return;
    char *base = reinterpret_cast<char *>(s->key_in);
    (void)hipFree(base);
    delete s;
    f->pop = nullptr;
}

// ------------------------------------------------------------------- drop
__global__ void k_pop_drop_stage(const int64_t *__restrict__ np_dev, int64_t max_particles,
                                 const double *__restrict__ x, const double *__restrict__ y,
                                 const double *__restrict__ z, const double *__restrict__ gx,
                                 const double *__restrict__ gy, const double *__restrict__ gz,
                                 const double *__restrict__ sg, const double *__restrict__ age,
                                 double *__restrict__ tmp, int64_t cap)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t np = *np_dev;
    if (np <= max_particles || k >= max_particles) return;
    const int64_t s = k + (np - max_particles);
    const double *src[8] = {x, y, z, gx, gy, gz, sg, age};
    for (int c = 0; c < 8; ++c) tmp[c * cap + k] = src[c][s];
}

__global__ void k_pop_drop_back(int64_t *__restrict__ np_dev, int64_t max_particles,
                                double *__restrict__ x, double *__restrict__ y,
                                double *__restrict__ z, double *__restrict__ gx,
                                double *__restrict__ gy, double *__restrict__ gz,
                                double *__restrict__ sg, double *__restrict__ age,
                                const double *__restrict__ tmp, int64_t cap)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t np = *np_dev;
    if (np <= max_particles || k >= max_particles) return;
    double *dst[8] = {x, y, z, gx, gy, gz, sg, age};
    for (int c = 0; c < 8; ++c) dst[c][k] = tmp[c * cap + k];
}

__global__ void k_pop_set_count(int64_t *__restrict__ np_dev, int64_t max_particles)
{
    if (threadIdx.x == 0 && *np_dev > max_particles) *np_dev = max_particles;
}

int vpm_drop_oldest(fv_vpm *f, int64_t max_particles)
{
    fv_ctx *ctx = f->ctx;
    FV_REQUIRE(ctx, max_particles >= 0, "negative particle limit");
    if (f->np_upper <= max_particles) return FV_OK;
    int rc = pop_ensure(f);
    if (rc) return rc;
    PopScratch *s = pop_of(f);
    const unsigned gb = (unsigned)ceil_div(max_particles > 0 ? max_particles : 1, kPopBlock);
    hipLaunchKernelGGL(k_pop_drop_stage, dim3(gb), dim3(kPopBlock), 0, ctx->stream, f->np_dev,
                       max_particles, f->x, f->y, f->z, f->gx, f->gy, f->gz, f->sig, f->age, s->tmp,
                       s->cap);
    FV_LAUNCH_CHECK(ctx);
    hipLaunchKernelGGL(k_pop_drop_back, dim3(gb), dim3(kPopBlock), 0, ctx->stream, f->np_dev,
                       max_particles, f->x, f->y, f->z, f->gx, f->gy, f->gz, f->sig, f->age, s->tmp,
                       s->cap);
    FV_LAUNCH_CHECK(ctx);
    hipLaunchKernelGGL(k_pop_set_count, dim3(1), dim3(64), 0, ctx->stream, f->np_dev, max_particles);
    FV_LAUNCH_CHECK(ctx);
    f->np_upper = max_particles;
    f->np_exact = false;
    f->state_epoch++;
    return FV_OK;
}

// ------------------------------------------------------------------ merge
#pragma clang fp contract(off)

__device__ __forceinline__ void pop_centre(const double *ref, int nref, double *c)
{
    c[0] = c[1] = c[2] = 0.0;
    for (int k = 0; k < nref; ++k)
        for (int q = 0; q < 3; ++q) c[q] += ref[3 * k + q];
    for (int q = 0; q < 3; ++q) c[q] /= (double)nref;
}

__global__ void k_pop_keys(const int64_t *__restrict__ np_dev, int64_t n, const double *__restrict__ x,
                           const double *__restrict__ y, const double *__restrict__ z,
                           const double *__restrict__ ref, int nref, int64_t protect,
                           double protect_dist, uint64_t *__restrict__ key,
                           int64_t *__restrict__ val, int32_t *__restrict__ pair_at,
                           int32_t *__restrict__ keep, int64_t *__restrict__ counters)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int64_t np = *np_dev;
    pair_at[k] = 0;
    keep[k] = k < np ? 1 : 0;
    val[k] = k;
    key[k] = kSentinel;
    if (k >= np || np < 2 * protect) return;
    double c[3];
    pop_centre(ref, nref, c);
    const double px = x[k], py = y[k], pz = z[k];
    const double dx = px - c[0], dy = py - c[1], dz = pz - c[2];
    const double d = sqrt(dx * dx + dy * dy + dz * dz);
    const bool cand = protect_dist > 0.0 ? d > protect_dist : k < np - protect;
    if (!cand) return;
    atomicAdd(reinterpret_cast<unsigned long long *>(counters), 1ull);
    const double cell = 0.1 + 0.15 * d;
    const double q[3] = {floor(px / cell), floor(py / cell), floor(pz / cell)};
    uint64_t packed = 0;
    for (int a = 0; a < 3; ++a) {
        if (!(q[a] > -(double)kKeyBias && q[a] < (double)kKeyBias)) {
            atomicAdd(reinterpret_cast<unsigned long long *>(counters + 2), 1ull);
            return;   // beyond the key range: never merged
        }
        packed = (packed << 21) | (uint64_t)((int64_t)q[a] + kKeyBias);
    }
    key[k] = packed;
}

// One thread per run of equal keys: the greedy scan of the sorted order.
__global__ void k_pop_pairs(int64_t n, const uint64_t *__restrict__ key,
                            const int64_t *__restrict__ val, const double *__restrict__ gx,
                            const double *__restrict__ gy, const double *__restrict__ gz,
                            double protect_dist, const int64_t *__restrict__ counters,
                            int32_t *__restrict__ pair_at)
{
    const int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= n) return;
    const uint64_t ka = key[a];
    if (ka == kSentinel || (a > 0 && key[a - 1] == ka)) return;
    if (protect_dist > 0.0 && counters[0] < 4) return;
    int64_t b = a;
    while (b + 1 < n && key[b + 1] == ka) {
        const int64_t i = val[b], j = val[b + 1];
        const double dot = gx[i] * gx[j] + gy[i] * gy[j] + gz[i] * gz[j];
        if (dot > 0.0) {
            pair_at[b] = 1;
            b += 2;
        } else {
            b += 1;
        }
    }
}

__device__ __forceinline__ void pop_vel(const double *R, const double *X, const double *A, double S,
                                        double *v)
{
    const double r[3] = {R[0] - X[0], R[1] - X[1], R[2] - X[2]};
    const double d = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]) + 1e-30;
    const double rho = d / S;
    const double t = rho * rho + 1.0;
    const double g = rho * rho * rho / (t * sqrt(t));
    const double c = g / (4.0 * M_PI * (d * d * d));
    v[0] = c * (A[1] * r[2] - A[2] * r[1]);
    v[1] = c * (A[2] * r[0] - A[0] * r[2]);
    v[2] = c * (A[0] * r[1] - A[1] * r[0]);
}

__global__ void k_pop_eval(int64_t n, const int32_t *__restrict__ pair_at,
                           const int64_t *__restrict__ val, double *__restrict__ x,
                           double *__restrict__ y, double *__restrict__ z,
                           double *__restrict__ gx, double *__restrict__ gy,
                           double *__restrict__ gz, double *__restrict__ sg,
                           const double *__restrict__ ref, int nref, double thr,
                           int32_t *__restrict__ keep, int64_t *__restrict__ counters)
{
    const int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= n || !pair_at[a]) return;
    const int64_t i = val[a], j = val[a + 1];
    const double a1[3] = {gx[i], gy[i], gz[i]}, a2[3] = {gx[j], gy[j], gz[j]};
    const double x1[3] = {x[i], y[i], z[i]}, x2[3] = {x[j], y[j], z[j]};
    const double w1 = sqrt(a1[0] * a1[0] + a1[1] * a1[1] + a1[2] * a1[2]);
    const double w2 = sqrt(a2[0] * a2[0] + a2[1] * a2[1] + a2[2] * a2[2]);
    double am[3], xm[3];
    for (int k = 0; k < 3; ++k) {
        am[k] = a1[k] + a2[k];
        xm[k] = (w1 * x1[k] + w2 * x2[k]) / (w1 + w2 + 1e-30);
    }
    const double s1 = sg[i], s2 = sg[j];
    const double sm = s1 > s2 ? s1 : s2;
    double err = 0.0;
    for (int r = 0; r < nref; ++r) {
        double v1[3], v2[3], v3[3];
        pop_vel(ref + 3 * r, x1, a1, s1, v1);
        pop_vel(ref + 3 * r, x2, a2, s2, v2);
        pop_vel(ref + 3 * r, xm, am, sm, v3);
        for (int k = 0; k < 3; ++k) {
            const double dv = fabs((v1[k] + v2[k]) - v3[k]);
            if (dv > err) err = dv;
        }
    }
    if (!(err < thr)) return;
    x[i] = xm[0]; y[i] = xm[1]; z[i] = xm[2];
    gx[i] = am[0]; gy[i] = am[1]; gz[i] = am[2];
    sg[i] = sm;
    keep[j] = 0;
    atomicAdd(reinterpret_cast<unsigned long long *>(counters + 1), 1ull);
}

#pragma clang fp contract(on)

__global__ void k_pop_compact(int64_t n, const int32_t *__restrict__ keep,
                              const int32_t *__restrict__ pos, const double *__restrict__ x,
                              const double *__restrict__ y, const double *__restrict__ z,
                              const double *__restrict__ gx, const double *__restrict__ gy,
                              const double *__restrict__ gz, const double *__restrict__ sg,
                              const double *__restrict__ age, double *__restrict__ tmp, int64_t cap)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n || !keep[k]) return;
    const int64_t w = pos[k];
    const double *src[8] = {x, y, z, gx, gy, gz, sg, age};
    for (int c = 0; c < 8; ++c) tmp[c * cap + w] = src[c][k];
}

__global__ void k_pop_compact_back(int64_t n, const int32_t *__restrict__ keep,
                                   const int32_t *__restrict__ pos, int64_t *__restrict__ np_dev,
                                   double *__restrict__ x, double *__restrict__ y,
                                   double *__restrict__ z, double *__restrict__ gx,
                                   double *__restrict__ gy, double *__restrict__ gz,
                                   double *__restrict__ sg, double *__restrict__ age,
                                   const double *__restrict__ tmp, int64_t cap)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t total = n > 0 ? (int64_t)pos[n - 1] + keep[n - 1] : 0;
    if (k == 0) *np_dev = total;
    if (k >= total) return;
    double *dst[8] = {x, y, z, gx, gy, gz, sg, age};
    for (int c = 0; c < 8; ++c) dst[c][k] = tmp[c * cap + k];
}

__global__ void k_pop_zero_candidates(int64_t *__restrict__ counters)
{
    if (threadIdx.x == 0) counters[0] = 0;
}

int vpm_merge(fv_vpm *f, const double *ref_dev, int nref, double eps, double vinf_norm,
              int64_t protect, double protect_dist)
{
    fv_ctx *ctx = f->ctx;
    FV_REQUIRE(ctx, nref >= 1 && nref <= kMaxRef, "reference point count %d outside 1..%d", nref,
               kMaxRef);
    FV_REQUIRE(ctx, eps >= 0.0 && vinf_norm >= 0.0 && protect >= 0 && protect_dist >= 0.0,
               "negative merge parameter");
    const int64_t n = f->np_upper;
    if (n < 2) return FV_OK;
    int rc = pop_ensure(f);
    if (rc) return rc;
    PopScratch *s = pop_of(f);
    const unsigned gb = (unsigned)ceil_div(n, kPopBlock);
    hipLaunchKernelGGL(k_pop_zero_candidates, dim3(1), dim3(64), 0, ctx->stream, s->counters);
    FV_LAUNCH_CHECK(ctx);
    hipLaunchKernelGGL(k_pop_keys, dim3(gb), dim3(kPopBlock), 0, ctx->stream, f->np_dev, n, f->x,
                       f->y, f->z, ref_dev, nref, protect, protect_dist, s->key_in, s->val_in,
                       s->pair_at, s->keep, s->counters);
    FV_LAUNCH_CHECK(ctx);
    size_t tb = s->cub_bytes;
    FV_HIP_CHECK(ctx, hipcub::DeviceRadixSort::SortPairs(s->cub, tb, s->key_in, s->key_out,
                                                         s->val_in, s->val_out, (int)n, 0, 64,
                                                         ctx->stream));
    hipLaunchKernelGGL(k_pop_pairs, dim3(gb), dim3(kPopBlock), 0, ctx->stream, n, s->key_out,
                       s->val_out, f->gx, f->gy, f->gz, protect_dist, s->counters, s->pair_at);
    FV_LAUNCH_CHECK(ctx);
    hipLaunchKernelGGL(k_pop_eval, dim3(gb), dim3(kPopBlock), 0, ctx->stream, n, s->pair_at,
                       s->val_out, f->x, f->y, f->z, f->gx, f->gy, f->gz, f->sig, ref_dev, nref,
                       eps * vinf_norm, s->keep, s->counters);
    FV_LAUNCH_CHECK(ctx);
    tb = s->cub_bytes;
    FV_HIP_CHECK(ctx, hipcub::DeviceScan::ExclusiveSum(s->cub, tb, s->keep, s->keep_pos, (int)n,
                                                       ctx->stream));
    hipLaunchKernelGGL(k_pop_compact, dim3(gb), dim3(kPopBlock), 0, ctx->stream, n, s->keep,
                       s->keep_pos, f->x, f->y, f->z, f->gx, f->gy, f->gz, f->sig, f->age, s->tmp,
                       s->cap);
    FV_LAUNCH_CHECK(ctx);
    hipLaunchKernelGGL(k_pop_compact_back, dim3(gb), dim3(kPopBlock), 0, ctx->stream, n, s->keep,
                       s->keep_pos, f->np_dev, f->x, f->y, f->z, f->gx, f->gy, f->gz, f->sig,
                       f->age, s->tmp, s->cap);
    FV_LAUNCH_CHECK(ctx);
    f->np_exact = false;
    f->state_epoch++;
    return FV_OK;
}

// Reference points of the hybrid driver: collocation points 0, s, 2s, ... (s =
// max(1, P / 9)), at most 9 (flap.py:426).
__global__ void k_pop_gather_ref(const double *__restrict__ colloc, int64_t stride, int nref,
                                 double *__restrict__ ref)
{
    const int t = threadIdx.x;
    if (t < 3 * nref) ref[t] = colloc[3 * (t / 3) * stride + t % 3];
}

int vpm_merge_colloc(fv_vpm *f, const double *colloc, int64_t P, double eps, double vinf_norm,
                     int64_t protect, double protect_dist)
{
    fv_ctx *ctx = f->ctx;
    FV_REQUIRE(ctx, P >= 1, "no collocation points");
    int rc = pop_ensure(f);
    if (rc) return rc;
    const int64_t stride = P / 9 > 1 ? P / 9 : 1;
    const int nref = (int)((P + stride - 1) / stride < 9 ? (P + stride - 1) / stride : 9);
    hipLaunchKernelGGL(k_pop_gather_ref, dim3(1), dim3(64), 0, ctx->stream, colloc, stride, nref,
                       pop_of(f)->ref);
    FV_LAUNCH_CHECK(ctx);
    return vpm_merge(f, pop_of(f)->ref, nref, eps, vinf_norm, protect, protect_dist);
}

int vpm_merge_stats(fv_vpm *f, int64_t *merged, int64_t *out_of_range)
{
    fv_ctx *ctx = f->ctx;
    int64_t c[4] = {0, 0, 0, 0};
    if (f->pop) {
        FV_HIP_CHECK(ctx, hipMemcpyAsync(c, pop_of(f)->counters, sizeof c, hipMemcpyDeviceToHost,
                                         ctx->stream));
        FV_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    }
    if (merged) *merged = c[1];
    if (out_of_range) *out_of_range = c[2];
    return FV_OK;
}

// ----------------------------------------------------------------- ledger
constexpr int kLedgerBlock = 256;

// Partial sums of alpha over fixed particle ranges (per block, wave order).
__global__ __launch_bounds__(kLedgerBlock) void k_ledger_part(
    const int64_t *__restrict__ np_dev, const double *__restrict__ gx,
    const double *__restrict__ gy, const double *__restrict__ gz, double *__restrict__ part)
{
    __shared__ double red[kLedgerBlock / 64][3];
    const int64_t np = *np_dev;
    const int64_t k = (int64_t)blockIdx.x * kLedgerBlock + threadIdx.x;
    double a[3] = {0.0, 0.0, 0.0};
    if (k < np) {
        a[0] = gx[k];
        a[1] = gy[k];
        a[2] = gz[k];
    }
    for (int c = 0; c < 3; ++c)
        for (int off = 32; off > 0; off >>= 1) a[c] += __shfl_xor(a[c], off, 64);
    if ((threadIdx.x & 63) == 0)
        for (int c = 0; c < 3; ++c) red[threadIdx.x >> 6][c] = a[c];
    __syncthreads();
    if (threadIdx.x == 0)
        for (int c = 0; c < 3; ++c) {
            double t = 0.0;
            for (int w = 0; w < kLedgerBlock / 64; ++w) t += red[w][c];
            part[3 * blockIdx.x + c] = t;
        }
}

// Ring legs: G (b - a) over Fl->Fr, Fr->Br, Br->Bl, Bl->Fl per ring.
__global__ void k_ledger_rings(int64_t nr, const double *__restrict__ br,
                               const double *__restrict__ fr, const double *__restrict__ fl,
                               const double *__restrict__ bl, const double *__restrict__ g,
                               double *__restrict__ part)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nr) return;
    const double *cc[4] = {fl + 3 * k, fr + 3 * k, br + 3 * k, bl + 3 * k};
    double t[3] = {0.0, 0.0, 0.0};
    for (int l = 0; l < 4; ++l)
        for (int q = 0; q < 3; ++q) t[q] = t[q] + g[k] * (cc[(l + 1) & 3][q] - cc[l][q]);
    for (int q = 0; q < 3; ++q) part[3 * k + q] = t[q];
}

__global__ void k_ledger_final(const double *__restrict__ part, int64_t nparts,
                               double *__restrict__ out)
{
    if (threadIdx.x != 0) return;
    double t[3] = {0.0, 0.0, 0.0};
    for (int64_t b = 0; b < nparts; ++b)
        for (int c = 0; c < 3; ++c) t[c] += part[3 * b + c];
    for (int c = 0; c < 3; ++c) out[c] = t[c];
}

// Enqueues the ledger into out3 (device): particles, then the rings (may be
// nr = 0).  Uses the population scratch.
int vpm_ledger(fv_vpm *f, int64_t nr, const double *br, const double *fr, const double *fl,
               const double *bl, const double *g, double *out3)
{
    fv_ctx *ctx = f->ctx;
    int rc = pop_ensure(f);
    if (rc) return rc;
    PopScratch *s = pop_of(f);
    const int64_t nb = ceil_div(f->np_upper > 0 ? f->np_upper : 1, kLedgerBlock);
    FV_REQUIRE(ctx, 3 * (nb + nr + 1) <= 8 * s->cap, "ledger: too many rings (%lld)",
               (long long)nr);
    double *part = s->tmp;
    hipLaunchKernelGGL(k_ledger_part, dim3((unsigned)nb), dim3(kLedgerBlock), 0, ctx->stream,
                       f->np_dev, f->gx, f->gy, f->gz, part);
    FV_LAUNCH_CHECK(ctx);
    if (nr > 0) {
        hipLaunchKernelGGL(k_ledger_rings, dim3((unsigned)ceil_div(nr, 128)), dim3(128), 0,
                           ctx->stream, nr, br, fr, fl, bl, g, part + 3 * nb);
        FV_LAUNCH_CHECK(ctx);
    }
    hipLaunchKernelGGL(k_ledger_final, dim3(1), dim3(64), 0, ctx->stream, part, nb + nr, out3);
    FV_LAUNCH_CHECK(ctx);
    return FV_OK;
}

}  // namespace fv

using namespace fv;

extern "C" {

int fv_vpm_drop_oldest(fv_vpm *f, int64_t max_particles)
{
    if (!f) return FV_E_ARG;
    return vpm_drop_oldest(f, max_particles);
}

int fv_vpm_merge_host(fv_vpm *f, const double *ref, int32_t nref, double eps, double vinf_norm,
                      int64_t protect, double protect_dist, int64_t *n_merged)
{
    if (!f || !ref) return FV_E_ARG;
    fv_ctx *ctx = f->ctx;
    FV_REQUIRE(ctx, nref >= 1 && nref <= kMaxRef, "reference point count %d outside 1..%d", nref,
               kMaxRef);
    int rc = pop_ensure(f);
    if (rc) return rc;
    PopScratch *s = pop_of(f);
    int64_t before = 0;
    if (n_merged) {
        rc = vpm_merge_stats(f, &before, nullptr);
        if (rc) return rc;
    }
    FV_HIP_CHECK(ctx, hipMemcpyAsync(s->ref, ref, sizeof(double) * 3 * nref,
                                     hipMemcpyHostToDevice, ctx->stream));
    rc = vpm_merge(f, s->ref, nref, eps, vinf_norm, protect, protect_dist);
    if (rc) return rc;
    if (n_merged) {
        int64_t after = 0;
        rc = vpm_merge_stats(f, &after, nullptr);
        if (rc) return rc;
        *n_merged = after - before;
    }
    return FV_OK;
}

int fv_vpm_merge_stats(fv_vpm *f, int64_t *merged, int64_t *out_of_range)
{
    if (!f) return FV_E_ARG;
    return vpm_merge_stats(f, merged, out_of_range);
}

int fv_vpm_ledger_host(fv_vpm *f, double *out)
{
    if (!f || !out) return FV_E_ARG;
    fv_ctx *ctx = f->ctx;
    int rc = pop_ensure(f);
    if (rc) return rc;
    double *d = pop_of(f)->ref;   // 3 doubles of the reference-point slot
    rc = vpm_ledger(f, 0, nullptr, nullptr, nullptr, nullptr, nullptr, d);
    if (rc) return rc;
    FV_HIP_CHECK(ctx, hipMemcpyAsync(out, d, 3 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    FV_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return FV_OK;
}

}  // extern "C"
