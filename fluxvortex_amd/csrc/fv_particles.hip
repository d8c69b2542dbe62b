// fv_particles.hip -- vortex particle sums (velocity, velocity gradient,
// stretching, vorticity), the device-resident particle field with its RK3 +
// rVPM + Pedrizzetti step, and panel -> particle conversion.
//
// Reference: kernel.py:69-190 (NumPy chunked sums), warp_vpm.py:19-173 (one
// thread per pair + atomics), particles.py:15-246 (host RK3 loop),
// tests/experiment_hybrid_panel_particle.py:145-204 (conversion).
#include <cmath>
#include <cstring>
#include <new>
#include <vector>

#include "fv_vpm.h"

// (n, 3) from three columns, defined with the host-API staging below
extern "C" __global__ void k_soa_to_aos(const double *__restrict__ a, const double *__restrict__ b,
                             const double *__restrict__ c, int64_t n, double *__restrict__ out);

namespace fv {

// ---------------------------------------------------------- RK3 stages
// Position stage, particles.py:121-144:  U = U_part + U_inf, |U| <= 50,
// rkX = a rkX + dt U, x += b rkX.  U_part is null for a prescribed wake.
__global__ void k_vpm_stage_pos(const int64_t *__restrict__ np_dev, double *__restrict__ x,
                                double *__restrict__ y, double *__restrict__ z,
                                const double *__restrict__ upart, double ux, double uy,
                                double uz, const double *__restrict__ ub,
                                const double *__restrict__ ui, double a, double b, double dt,
                                double *__restrict__ rkX)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= *np_dev) return;
    double u0 = ux, u1 = uy, u2 = uz;
    if (ub) {
        // the callables' fields: (U_part + U_bound) + U_inf, :133
        const double p0 = upart ? upart[3 * i] : 0.0, p1 = upart ? upart[3 * i + 1] : 0.0,
                     p2 = upart ? upart[3 * i + 2] : 0.0;
        u0 = (p0 + ub[3 * i]) + ui[3 * i];
        u1 = (p1 + ub[3 * i + 1]) + ui[3 * i + 1];
        u2 = (p2 + ub[3 * i + 2]) + ui[3 * i + 2];
    } else if (upart) {
        u0 = (upart[3 * i] + 0.0) + ux;
        u1 = (upart[3 * i + 1] + 0.0) + uy;
        u2 = (upart[3 * i + 2] + 0.0) + uz;
    }
    const double sp = sqrt(u0 * u0 + u1 * u1 + u2 * u2);
    if (sp > 50.0) {
        const double f = 50.0 / sp;
        u0 *= f;
        u1 *= f;
        u2 *= f;
    }
    double *r = rkX + 3 * i;
    r[0] = a * r[0] + dt * u0;
    r[1] = a * r[1] + dt * u1;
    r[2] = a * r[2] + dt * u2;
    x[i] += b * r[0];
    y[i] += b * r[1];
    z[i] += b * r[2];
}

// Stretching stage, particles.py:146-183, with S = J . Gamma precomputed.
__global__ void k_vpm_stage_gamma(const int64_t *__restrict__ np_dev,
                                  const double *__restrict__ S, double *__restrict__ gx,
                                  double *__restrict__ gy, double *__restrict__ gz,
                                  double *__restrict__ sig, double a, double b, double dt,
                                  double nu, double *__restrict__ rkG, double *__restrict__ rks)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= *np_dev) return;
    double g[3] = {gx[i], gy[i], gz[i]};
    const double s0 = S[3 * i], s1 = S[3 * i + 1], s2 = S[3 * i + 2];
    const double gsq = g[0] * g[0] + g[1] * g[1] + g[2] * g[2];
    const double sdg = s0 * g[0] + s1 * g[1] + s2 * g[2];
    double Z = 0.0;
    if (gsq > 1e-20) Z = 0.2 * sdg / gsq;
    const double zmax = 10.0 / (dt > 1e-10 ? dt : 1e-10);
    Z = Z < -zmax ? -zmax : (Z > zmax ? zmax : Z);
    const double sg = sig[i];
    double ds = -sg * Z;
    if (nu > 0) ds += nu / (sg > 1e-10 ? sg : 1e-10);
    const double dG[3] = {s0 - 3.0 * Z * g[0], s1 - 3.0 * Z * g[1], s2 - 3.0 * Z * g[2]};
    double *rg = rkG + 3 * i;
    for (int c = 0; c < 3; ++c) {
        rg[c] = a * rg[c] + dt * dG[c];
        g[c] += b * rg[c];
    }
    rks[i] = a * rks[i] + dt * ds;
    double sn = sg + b * rks[i];
    const double g2 = g[0] * g[0] + g[1] * g[1] + g[2] * g[2];
    if (g2 > 1e4) {
        const double sc = sqrt(1e4 / g2);
        g[0] *= sc;
        g[1] *= sc;
        g[2] *= sc;
    }
    sn = sn < 1e-6 ? 1e-6 : (sn > 5.0 ? 5.0 : sn);
    gx[i] = g[0];
    gy[i] = g[1];
    gz[i] = g[2];
    sig[i] = sn;
}

// particles.py:185-196: non-finite position or circulation -> Gamma = 0; age += dt.
// Thread 0 also counts this advect's direct pair sums: passes x n^2 pairs
// into pairs_dev[0] (exact count, no host synchronisation).
__global__ void k_vpm_finish(const int64_t *__restrict__ np_dev, const double *__restrict__ x,
                             const double *__restrict__ y, const double *__restrict__ z,
                             double *__restrict__ gx, double *__restrict__ gy,
                             double *__restrict__ gz, double *__restrict__ age, double dt,
                             double passes, double *__restrict__ pairs_dev)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t n = *np_dev;
    if (i == 0) pairs_dev[0] += passes * (double)n * (double)n;
    if (i >= n) return;
    const bool bad = !isfinite(x[i]) || !isfinite(y[i]) || !isfinite(z[i]) ||
                     !isfinite(gx[i]) || !isfinite(gy[i]) || !isfinite(gz[i]);
    if (bad) {
        gx[i] = 0.0;
        gy[i] = 0.0;
        gz[i] = 0.0;
    }
    age[i] += dt;
}

// Pedrizzetti relaxation, particles.py:212-246, with the vorticity
// omega_i = eps_ijk J_kj of the Gaussian-erf field precomputed in W: three
// components per particle, or six (split = true: the clamped pairs' sum in
// 3..5, added last).
__global__ void k_vpm_relax(const int64_t *__restrict__ np_dev, const double *__restrict__ W,
                            int split, double *__restrict__ gx, double *__restrict__ gy,
                            double *__restrict__ gz, double rlxf)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= *np_dev) return;
    double g[3] = {gx[i], gy[i], gz[i]};
    const double gn = sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]);
    const double *Wi = W + (split ? 6 : 3) * i;
    const double w[3] = {split ? Wi[0] + Wi[3] : Wi[0], split ? Wi[1] + Wi[4] : Wi[1],
                         split ? Wi[2] + Wi[5] : Wi[2]};
    const double wn = sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
    if (!(gn > 1e-15 && wn > 1e-12)) return;
    const double wh[3] = {w[0] / wn, w[1] / wn, w[2] / wn};
    for (int c = 0; c < 3; ++c) g[c] = (1.0 - rlxf) * g[c] + rlxf * gn * wh[c];
    const double dot = g[0] * wh[0] + g[1] * wh[1] + g[2] * wh[2];
    double b2 = 1.0 - 2.0 * (1.0 - rlxf) * rlxf * (1.0 - dot / (gn > 1e-12 ? gn : 1e-12));
    b2 = b2 < 0.01 ? 0.01 : b2;
    const double sb = sqrt(b2);
    gx[i] = g[0] / sb;
    gy[i] = g[1] / sb;
    gz[i] = g[2] / sb;
}

// Conversion of nr rings into particles, one workgroup, rings in order.
// A ring with |gamma| >= 1e-15 yields its front, left, back and right leg
// particles (alpha = gamma (b - a), kept when |alpha|^2 > 1e-30) and its gamma
// is set to 0 (experiment_hybrid_panel_particle.py:173-200).
constexpr int kConvertBlock = 1024;

__global__ __launch_bounds__(kConvertBlock) void k_convert_rings(
    int64_t nr, const double *__restrict__ br, const double *__restrict__ fr,
    const double *__restrict__ fl, const double *__restrict__ bl, double *__restrict__ gamma,
    double sigma, int64_t cap, int64_t *__restrict__ np_dev, int32_t *__restrict__ overflow,
    double *__restrict__ x, double *__restrict__ y, double *__restrict__ z,
    double *__restrict__ gx, double *__restrict__ gy, double *__restrict__ gz,
    double *__restrict__ sg, double *__restrict__ age)
{
    __shared__ int cnt[kConvertBlock];
    __shared__ int64_t start_s;   // particle count before the current chunk
    __shared__ int chunk_total_s;
    __shared__ int ovf_s;
    if (threadIdx.x == 0) {
        start_s = *np_dev;
        ovf_s = 0;
    }
    __syncthreads();
    for (int64_t r0 = 0; r0 < nr; r0 += kConvertBlock) {
        const int64_t r = r0 + threadIdx.x;
        double ps[4][3], al[4][3];
        int keep = 0;   // bits 0-3: kept legs, bit 4: ring converted
        if (r < nr) {
            const double G = gamma[r];
            if (fabs(G) >= 1e-15) {
                const double *FR = fr + 3 * r, *FL = fl + 3 * r, *BL = bl + 3 * r,
                             *BR = br + 3 * r;
                const double *ends[4][2] = {{FR, FL}, {FL, BL}, {BL, BR}, {BR, FR}};
                for (int l = 0; l < 4; ++l) {
                    for (int c = 0; c < 3; ++c) {
                        ps[l][c] = 0.5 * (ends[l][0][c] + ends[l][1][c]);
                        al[l][c] = G * (ends[l][1][c] - ends[l][0][c]);
                    }
                    if (al[l][0] * al[l][0] + al[l][1] * al[l][1] + al[l][2] * al[l][2] > 1e-30)
                        keep |= 1 << l;
                }
                keep |= 16;
            }
        }
        cnt[threadIdx.x] = __popc(keep & 15);
        __syncthreads();
        if (threadIdx.x == 0) {
            int run = 0;
            for (int k = 0; k < kConvertBlock; ++k) {
                const int c = cnt[k];
                cnt[k] = run;
                run += c;
            }
            chunk_total_s = run;
            if (start_s + run > cap) ovf_s = 1;
        }
        __syncthreads();
        if (keep & 16) {
            if (!ovf_s) {
                int64_t k = start_s + cnt[threadIdx.x];
                for (int l = 0; l < 4; ++l) {
                    if (!(keep & (1 << l))) continue;
                    x[k] = ps[l][0];
                    y[k] = ps[l][1];
                    z[k] = ps[l][2];
                    gx[k] = al[l][0];
                    gy[k] = al[l][1];
                    gz[k] = al[l][2];
                    sg[k] = sigma;
                    age[k] = 0.0;
                    ++k;
                }
            }
            gamma[r] = 0.0;   // zeroed before the capacity check, as the reference
        }
        __syncthreads();
        if (threadIdx.x == 0 && !ovf_s) start_s += chunk_total_s;
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        if (ovf_s)
            *overflow = 1;
        else
            *np_dev = start_s;
    }
}

int vpm_sync_count(fv_vpm *f)
{
    fv_ctx *ctx = f->ctx;
    int64_t n = 0;
    FV_HIP_CHECK(ctx, hipMemcpyAsync(&n, f->np_dev, sizeof n, hipMemcpyDeviceToHost, ctx->stream));
    FV_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    f->np_host = n;
    f->np_upper = n;
    f->np_exact = true;
    return FV_OK;
}

int vpm_convert_rings(fv_vpm *f, int64_t nr, const double *br, const double *fr,
                      const double *fl, const double *bl, double *gamma, double sigma)
{
    fv_ctx *ctx = f->ctx;
    if (nr <= 0) return FV_OK;
    TimedLaunch tl(ctx, "convert");
    hipLaunchKernelGGL(k_convert_rings, dim3(1), dim3(kConvertBlock), 0, ctx->stream, nr, br,
                       fr, fl, bl, gamma, sigma, f->cap, f->np_dev, f->overflow_dev, f->x, f->y,
                       f->z, f->gx, f->gy, f->gz, f->sig, f->age);
    FV_LAUNCH_CHECK(ctx);
    tl.done();
    f->np_exact = false;
    f->state_epoch++;
    f->np_upper += 4 * nr;
    if (f->np_upper > f->cap) f->np_upper = f->cap;
    return FV_OK;
}

static PartSrcArgs field_src(const fv_vpm *f, int kernel)
{
    PartSrcArgs s;
    s.x = f->x; s.y = f->y; s.z = f->z;
    s.ax = f->gx; s.ay = f->gy; s.az = f->gz;
    s.sig = f->sig;
    s.stride = 1;
    s.kernel = kernel;
    return s;
}

static PartTgtArgs field_tgt(const fv_vpm *f)
{
    PartTgtArgs t;
    t.x = f->x; t.y = f->y; t.z = f->z;
    t.ax = f->gx; t.ay = f->gy; t.az = f->gz;
    t.stride = 1;
    return t;
}

// pairs_dev[1] += targets x the device particle count (one thread)
__global__ void k_count_induce(const int64_t *__restrict__ np_dev, double targets,
                               double *__restrict__ pairs_dev)
{
    pairs_dev[1] += targets * (double)*np_dev;
}

int vpm_induce(fv_vpm *f, const double *targets, int64_t n, int precision, double *out,
               int pslot)
{
    PartTgtArgs t;
    t.x = targets; t.y = targets + 1; t.z = targets + 2;
    t.ax = t.ay = t.az = nullptr;
    t.stride = 3;
    const int64_t nb = launch_count(f);
    int rc = part_sum(f->ctx, f->kernel, precision, 0, field_src(f, f->kernel), t, n, nb,
                      nullptr, f->np_dev, out, pslot, "particle_induce");
    if (rc || n <= 0 || nb <= 0) return rc;
    hipLaunchKernelGGL(k_count_induce, dim3(1), dim3(1), 0, f->ctx->stream, f->np_dev,
                       (double)n, f->pairs_dev);
    FV_LAUNCH_CHECK(f->ctx);
    return FV_OK;
}

// The callables' fields at the current (stage) positions into the staging
// buffer: positions to the host, the callback, u_bound then u_inf back
// (f->stage[0, 3n) and [3n, 6n)); synchronous.
static int stage_fields(fv_vpm *f, fv_stage_velocity_fn ext, void *user, std::vector<double> &h)
{
    fv_ctx *ctx = f->ctx;
    const int64_t n = f->np_host;
    h.assign((size_t)9 * n, 0.0);
    hipLaunchKernelGGL(k_soa_to_aos, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, ctx->stream,
                       f->x, f->y, f->z, n, f->stage);
    FV_LAUNCH_CHECK(ctx);
    FV_HIP_CHECK(ctx, hipMemcpyAsync(h.data(), f->stage, sizeof(double) * 3 * n,
                                     hipMemcpyDeviceToHost, ctx->stream));
    FV_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    const int r = ext(user, n, h.data(), h.data() + 3 * n, h.data() + 6 * n);
    if (r != 0) return set_error(ctx, FV_E_ARG, "stage velocity callback returned %d", r);
    FV_HIP_CHECK(ctx, hipMemcpyAsync(f->stage, h.data() + 3 * n, sizeof(double) * 6 * n,
                                     hipMemcpyHostToDevice, ctx->stream));
    FV_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return FV_OK;
}

int vpm_advect(fv_vpm *f, double dt, const double *uinf, int stretch, int free_wake,
               fv_stage_velocity_fn ext, void *user)
{
    fv_ctx *ctx = f->ctx;
    if (ext && !f->np_exact) {
        const int rc = vpm_sync_count(f);
        if (rc) return rc;
    }
    const int64_t nu = launch_count(f);
    if (nu <= 0) return FV_OK;
    static const double kZero[3] = {0.0, 0.0, 0.0};
    if (ext) uinf = kZero;
    std::vector<double> hbuf;
    static const double ra[3] = {0.0, -5.0 / 9.0, -153.0 / 128.0};   // particles.py:114
    static const double rb[3] = {1.0 / 3.0, 15.0 / 16.0, 8.0 / 15.0};
    int rc = FV_OK;
    if (f->closure) {
        rc = closure_step(f);
        if (rc) return rc;
    }
    const bool clo = closure_active(f);
    rc = fill(ctx, f->rkX, 3 * nu, 0.0);
    if (!rc && stretch) rc = fill(ctx, f->rkG, 3 * nu, 0.0);
    if (!rc && stretch) rc = fill(ctx, f->rks, nu, 0.0);
    if (rc) return rc;
    const unsigned gb = (unsigned)ceil_div(nu, 256);
    const PartSrcArgs src = field_src(f, f->kernel);
    const PartTgtArgs tgt = field_tgt(f);
    // self sum of one mode, by the closure when it is in use, else directly
    // direct self-sum passes of this advect (counted on the device)
    const double passes = clo ? 0.0
                              : (free_wake ? 3.0 : 0.0) + (stretch ? 3.0 : 0.0) +
                                    (f->rlxf > 0 ? 1.0 : 0.0);
    auto self_sum = [&](int mode, int kernel) -> int {
        if (clo) return closure_self_sum(f, mode, kernel, f->sum);
        return part_sum(ctx, kernel, f->precision, mode,
                        kernel == f->kernel ? src : field_src(f, kernel), tgt, nu, nu, f->np_dev,
                        f->np_dev, f->sum, S_TMP0);
    };
    for (int st = 0; st < 3; ++st) {
        const double *upart = nullptr;
        if (free_wake) {
            rc = self_sum(0, f->kernel);
            if (rc) return rc;
            upart = f->sum;
        }
        const double *ub = nullptr, *ui = nullptr;
        if (ext) {
            rc = stage_fields(f, ext, user, hbuf);
            if (rc) return rc;
            ub = f->stage;
            ui = f->stage + 3 * f->np_host;
        }
        hipLaunchKernelGGL(k_vpm_stage_pos, dim3(gb), dim3(256), 0, ctx->stream, f->np_dev, f->x,
                           f->y, f->z, upart, uinf[0], uinf[1], uinf[2], ub, ui, ra[st], rb[st],
                           dt, f->rkX);
        FV_LAUNCH_CHECK(ctx);
        f->state_epoch++;
        if (stretch) {
            rc = self_sum(2, f->kernel);
            if (rc) return rc;
            hipLaunchKernelGGL(k_vpm_stage_gamma, dim3(gb), dim3(256), 0, ctx->stream, f->np_dev,
                               f->sum, f->gx, f->gy, f->gz, f->sig, ra[st], rb[st], dt, f->nu,
                               f->rkG, f->rks);
            FV_LAUNCH_CHECK(ctx);
            f->state_epoch++;
        }
    }
    hipLaunchKernelGGL(k_vpm_finish, dim3(gb), dim3(256), 0, ctx->stream, f->np_dev, f->x, f->y,
                       f->z, f->gx, f->gy, f->gz, f->age, dt, passes, f->pairs_dev);
    FV_LAUNCH_CHECK(ctx);
    f->state_epoch++;
    if (f->rlxf > 0) {
        // the relaxation always uses the Gaussian-erf field (particles.py:219);
        // under the reference CPU pair rule its Jacobian keeps the self pair
        // and coincident particles with r = 1e-12 (kernel.py:156), mode 4
        const int mode = f->pair_rule == FV_PAIRS_CLAMP ? 4 : 3;
        rc = self_sum(mode, FV_KERNEL_GAUSSIAN_ERF);
        if (rc) return rc;
        // the direct mode-4 sum keeps the clamped pairs apart (6 components);
        // the closure's near field has already added them
        const int split = (mode == 4 && !clo) ? 1 : 0;
        hipLaunchKernelGGL(k_vpm_relax, dim3(gb), dim3(256), 0, ctx->stream, f->np_dev, f->sum,
                           split, f->gx, f->gy, f->gz, f->rlxf);
        FV_LAUNCH_CHECK(ctx);
        f->state_epoch++;
    }
    return FV_OK;
}

}  // namespace fv

using namespace fv;

static int check_particle_args(fv_ctx *ctx, const double *targets, int64_t t, const double *pos,
                               const double *alpha, const double *sigma, int64_t n,
                               int32_t kernel, int32_t precision, const void *out)
{
    if (!ctx) return FV_E_ARG;
    FV_REQUIRE(ctx, t >= 0 && n >= 0, "negative size");
    FV_REQUIRE(ctx, kernel == FV_KERNEL_GAUSSIAN_ERF || kernel == FV_KERNEL_WINCKELMANS,
               "unknown kernel %d", kernel);
    FV_REQUIRE(ctx, precision == FV_PREC_FP64 || precision == FV_PREC_MIXED,
               "unknown precision %d", precision);
    FV_REQUIRE(ctx, t == 0 || (targets && out), "null targets/out");
    FV_REQUIRE(ctx, n == 0 || (pos && alpha && sigma), "null particle arrays");
    return FV_OK;
}

static int particle_host(fv_ctx *ctx, int mode, const double *targets, int64_t t,
                         const double *pos, const double *alpha, const double *sigma, int64_t n,
                         int32_t kernel, int32_t precision, double *out)
{
    int rc = check_particle_args(ctx, targets, t, pos, alpha, sigma, n, kernel, precision, out);
    if (rc) return rc;
    const int comp = mode == 1 ? 9 : 3;
    if (t == 0) return FV_OK;
    if (n == 0) {
        memset(out, 0, sizeof(double) * comp * t);
        return FV_OK;
    }
    const double *dt = upload(ctx, S_PTGT, targets, 3 * t);
    const double *dp = upload(ctx, S_PPOS, pos, 3 * n);
    const double *da = upload(ctx, S_PALPHA, alpha, 3 * n);
    const double *ds = upload(ctx, S_PSIG, sigma, n);
    double *dout = static_cast<double *>(workspace(ctx, S_OUT, sizeof(double) * comp * t));
    if (!dt || !dp || !da || !ds || !dout) return FV_E_HIP;
    PartSrcArgs s;
    s.x = dp; s.y = dp + 1; s.z = dp + 2;
    s.ax = da; s.ay = da + 1; s.az = da + 2;
    s.sig = ds;
    s.stride = 3;
    s.kernel = kernel;
    PartTgtArgs tg;
    tg.x = dt; tg.y = dt + 1; tg.z = dt + 2;
    tg.ax = tg.ay = tg.az = nullptr;
    tg.stride = 3;
    rc = part_sum(ctx, kernel, precision, mode, s, tg, t, n, nullptr, nullptr, dout, S_PARTIAL);
    if (rc) return rc;
    return fv_memcpy_d2h(ctx, out, dout, sizeof(double) * comp * t);
}

extern "C" {

int fv_particle_velocity_host(fv_ctx *ctx, const double *targets, int64_t t,
                              const double *pos, const double *alpha, const double *sigma,
                              int64_t n, int32_t kernel, int32_t precision, double *out)
{
    return particle_host(ctx, 0, targets, t, pos, alpha, sigma, n, kernel, precision, out);
}

int fv_particle_jacobian_host(fv_ctx *ctx, const double *targets, int64_t t,
                              const double *pos, const double *alpha, const double *sigma,
                              int64_t n, int32_t kernel, int32_t precision, double *J)
{
    return particle_host(ctx, 1, targets, t, pos, alpha, sigma, n, kernel, precision, J);
}

int fv_vpm_create(fv_ctx *ctx, int64_t max_particles, double nu, double rlxf, int32_t kernel,
                  int32_t precision, fv_vpm **out)
{
    if (!ctx || !out) return FV_E_ARG;
    *out = nullptr;
    FV_REQUIRE(ctx, max_particles > 0, "max_particles must be positive");
    FV_REQUIRE(ctx, kernel == FV_KERNEL_GAUSSIAN_ERF || kernel == FV_KERNEL_WINCKELMANS,
               "unknown kernel %d", kernel);
    FV_REQUIRE(ctx, precision == FV_PREC_FP64 || precision == FV_PREC_MIXED,
               "unknown precision %d", precision);
    fv_vpm *f = new (std::nothrow) fv_vpm();
    if (!f) return set_error(ctx, FV_E_HIP, "out of host memory");
    f->ctx = ctx;
    f->cap = max_particles;
    f->nu = nu;
    f->rlxf = rlxf;
    f->kernel = kernel;
    f->precision = precision;
    const size_t cap = (size_t)max_particles;
    // 8 state + 3+3+1 RK + 9 sums + 9 staging doubles per particle, plus counters
    const size_t doubles = cap * (8 + 7 + 9 + 9);
    const size_t bytes = doubles * sizeof(double) + 256;
    hipError_t e = hipMalloc(&f->block, bytes);
    if (e != hipSuccess) {
        delete f;
        return set_error(ctx, FV_E_HIP, "hipMalloc(%zu) for particle field failed: %s", bytes,
                         hipGetErrorString(e));
    }
    double *p = static_cast<double *>(f->block);
    f->x = p; p += cap;
    f->y = p; p += cap;
    f->z = p; p += cap;
    f->gx = p; p += cap;
    f->gy = p; p += cap;
    f->gz = p; p += cap;
    f->sig = p; p += cap;
    f->age = p; p += cap;
    f->rkX = p; p += 3 * cap;
    f->rkG = p; p += 3 * cap;
    f->rks = p; p += cap;
    f->sum = p; p += 9 * cap;
    f->stage = p; p += 9 * cap;
    f->np_dev = reinterpret_cast<int64_t *>(p);
    f->overflow_dev = reinterpret_cast<int32_t *>(p + 1);
    f->pairs_dev = p + 2;
    const int64_t zero[4] = {0, 0, 0, 0};
    e = hipMemcpyAsync(f->np_dev, zero, sizeof zero, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
        (void)hipFree(f->block);
        delete f;
        return set_error(ctx, FV_E_HIP, "particle field init failed: %s", hipGetErrorString(e));
    }
    *out = f;
    return FV_OK;
}

void fv_vpm_destroy(fv_vpm *f)
{
    if (!f) return;
    closure_destroy(f->closure);
    f->closure = nullptr;
    (void)hipStreamSynchronize(f->ctx->stream);
    pop_destroy(f);
    (void)hipFree(f->block);
    delete f;
}

int fv_vpm_set_pair_rule(fv_vpm *f, int32_t rule)
{
    if (!f) return FV_E_ARG;
    FV_REQUIRE(f->ctx, rule == FV_PAIRS_SKIP_CLOSE || rule == FV_PAIRS_CLAMP,
               "unknown pair rule %d", rule);
    f->pair_rule = rule;
    return FV_OK;
}

int fv_vpm_count(fv_vpm *f, int64_t *np)
{
    if (!f || !np) return FV_E_ARG;
    if (!f->np_exact) {
        int rc = vpm_sync_count(f);
        if (rc) return rc;
    }
    *np = f->np_host;
    return FV_OK;
}

// AoS host (n,3) -> SoA device columns through the staging buffer.
__global__ void k_aos_to_soa(const double *__restrict__ in, int64_t n, double *__restrict__ a,
                             double *__restrict__ b, double *__restrict__ c)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    a[i] = in[3 * i];
    b[i] = in[3 * i + 1];
    c[i] = in[3 * i + 2];
}

__global__ void k_soa_to_aos(const double *__restrict__ a, const double *__restrict__ b,
                             const double *__restrict__ c, int64_t n, double *__restrict__ out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[3 * i] = a[i];
    out[3 * i + 1] = b[i];
    out[3 * i + 2] = c[i];
}

static int put_columns(fv_vpm *f, int64_t off, int64_t n, const double *pos, const double *gamma,
                       const double *sigma, const double *age)
{
    fv_ctx *ctx = f->ctx;
    const unsigned g = (unsigned)ceil_div(n, 256);
    if (pos) {
        FV_HIP_CHECK(ctx, hipMemcpyAsync(f->stage, pos, sizeof(double) * 3 * n,
                                         hipMemcpyHostToDevice, ctx->stream));
        hipLaunchKernelGGL(k_aos_to_soa, dim3(g), dim3(256), 0, ctx->stream, f->stage, n,
                           f->x + off, f->y + off, f->z + off);
        FV_LAUNCH_CHECK(ctx);
        FV_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    }
    if (gamma) {
        FV_HIP_CHECK(ctx, hipMemcpyAsync(f->stage, gamma, sizeof(double) * 3 * n,
                                         hipMemcpyHostToDevice, ctx->stream));
        hipLaunchKernelGGL(k_aos_to_soa, dim3(g), dim3(256), 0, ctx->stream, f->stage, n,
                           f->gx + off, f->gy + off, f->gz + off);
        FV_LAUNCH_CHECK(ctx);
        FV_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    }
    if (sigma)
        FV_HIP_CHECK(ctx, hipMemcpyAsync(f->sig + off, sigma, sizeof(double) * n,
                                         hipMemcpyHostToDevice, ctx->stream));
    if (age)
        FV_HIP_CHECK(ctx, hipMemcpyAsync(f->age + off, age, sizeof(double) * n,
                                         hipMemcpyHostToDevice, ctx->stream));
    FV_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return FV_OK;
}

int fv_vpm_add_host(fv_vpm *f, const double *pos, const double *gamma, const double *sigma,
                    int64_t n)
{
    if (!f) return FV_E_ARG;
    fv_ctx *ctx = f->ctx;
    FV_REQUIRE(ctx, n >= 0, "negative particle count");
    if (n == 0) return FV_OK;
    FV_REQUIRE(ctx, pos && gamma && sigma, "null particle arrays");
    int rc = f->np_exact ? FV_OK : vpm_sync_count(f);
    if (rc) return rc;
    if (f->np_host + n > f->cap)
        return set_error(ctx, FV_E_CAP, "Particle field would overflow (%lld+%lld > %lld)",
                         (long long)f->np_host, (long long)n, (long long)f->cap);
    rc = put_columns(f, f->np_host, n, pos, gamma, sigma, nullptr);
    if (rc) return rc;
    rc = fill(ctx, f->age + f->np_host, n, 0.0);
    if (rc) return rc;
    f->np_host += n;
    f->np_upper = f->np_host;
    f->state_epoch++;
    FV_HIP_CHECK(ctx, hipMemcpyAsync(f->np_dev, &f->np_host, sizeof(int64_t),
                                     hipMemcpyHostToDevice, ctx->stream));
    FV_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return FV_OK;
}

int fv_vpm_write_host(fv_vpm *f, int64_t np, const double *pos, const double *gamma,
                      const double *sigma, const double *age)
{
    if (!f) return FV_E_ARG;
    fv_ctx *ctx = f->ctx;
    FV_REQUIRE(ctx, np >= 0 && np <= f->cap, "particle count %lld outside 0..%lld",
               (long long)np, (long long)f->cap);
    FV_REQUIRE(ctx, np == 0 || (pos && gamma && sigma && age), "null particle arrays");
    int rc = np ? put_columns(f, 0, np, pos, gamma, sigma, age) : FV_OK;
    if (rc) return rc;
    f->np_host = np;
    f->np_upper = np;
    f->np_exact = true;
    f->state_epoch++;
    FV_HIP_CHECK(ctx, hipMemcpyAsync(f->np_dev, &f->np_host, sizeof(int64_t),
                                     hipMemcpyHostToDevice, ctx->stream));
    FV_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return FV_OK;
}

int fv_vpm_read_host(fv_vpm *f, double *pos, double *gamma, double *sigma, double *age)
{
    if (!f) return FV_E_ARG;
    fv_ctx *ctx = f->ctx;
    int rc = f->np_exact ? FV_OK : vpm_sync_count(f);
    if (rc) return rc;
    const int64_t n = f->np_host;
    if (n == 0) return FV_OK;
    const unsigned g = (unsigned)ceil_div(n, 256);
    if (pos) {
        hipLaunchKernelGGL(k_soa_to_aos, dim3(g), dim3(256), 0, ctx->stream, f->x, f->y, f->z, n,
                           f->stage);
        FV_LAUNCH_CHECK(ctx);
        FV_HIP_CHECK(ctx, hipMemcpyAsync(pos, f->stage, sizeof(double) * 3 * n,
                                         hipMemcpyDeviceToHost, ctx->stream));
        FV_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    }
    if (gamma) {
        hipLaunchKernelGGL(k_soa_to_aos, dim3(g), dim3(256), 0, ctx->stream, f->gx, f->gy, f->gz,
                           n, f->stage);
        FV_LAUNCH_CHECK(ctx);
        FV_HIP_CHECK(ctx, hipMemcpyAsync(gamma, f->stage, sizeof(double) * 3 * n,
                                         hipMemcpyDeviceToHost, ctx->stream));
        FV_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    }
    if (sigma)
        FV_HIP_CHECK(ctx, hipMemcpyAsync(sigma, f->sig, sizeof(double) * n, hipMemcpyDeviceToHost,
                                         ctx->stream));
    if (age)
        FV_HIP_CHECK(ctx, hipMemcpyAsync(age, f->age, sizeof(double) * n, hipMemcpyDeviceToHost,
                                         ctx->stream));
    FV_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return FV_OK;
}

int fv_vpm_induce_host(fv_vpm *f, const double *targets, int64_t t, double *out)
{
    if (!f) return FV_E_ARG;
    fv_ctx *ctx = f->ctx;
    FV_REQUIRE(ctx, t >= 0, "negative target count");
    if (t == 0) return FV_OK;
    FV_REQUIRE(ctx, targets && out, "null targets/out");
    const double *dt = upload(ctx, S_PTGT, targets, 3 * t);
    double *dout = static_cast<double *>(workspace(ctx, S_OUT, sizeof(double) * 3 * t));
    if (!dt || !dout) return FV_E_HIP;
    int rc = vpm_induce(f, dt, t, f->precision, dout, S_PARTIAL);
    if (rc) return rc;
    return fv_memcpy_d2h(ctx, out, dout, sizeof(double) * 3 * t);
}

int fv_vpm_advect_rk3(fv_vpm *f, double dt, const double *uinf, int32_t stretch,
                      int32_t free_wake)
{
    if (!f) return FV_E_ARG;
    fv_ctx *ctx = f->ctx;
    FV_REQUIRE(ctx, uinf, "null uinf");
    int rc = vpm_advect(f, dt, uinf, stretch, free_wake);
    if (rc) return rc;
    FV_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return FV_OK;
}

int fv_vpm_advect_rk3_ext(fv_vpm *f, double dt, fv_stage_velocity_fn fn, void *user,
                          int32_t stretch, int32_t free_wake)
{
    if (!f) return FV_E_ARG;
    fv_ctx *ctx = f->ctx;
    FV_REQUIRE(ctx, fn, "null stage velocity callback");
    int rc = vpm_advect(f, dt, nullptr, stretch, free_wake, fn, user);
    if (rc) return rc;
    FV_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return FV_OK;
}

int fv_convert_rings_host(fv_vpm *f, int64_t nr, const double *br, const double *fr,
                          const double *fl, const double *bl, double *gamma, double sigma,
                          int64_t *added)
{
    if (!f) return FV_E_ARG;
    fv_ctx *ctx = f->ctx;
    FV_REQUIRE(ctx, nr >= 0, "negative ring count");
    if (added) *added = 0;
    if (nr == 0) return FV_OK;
    FV_REQUIRE(ctx, br && fr && fl && bl && gamma, "null ring arrays");
    int rc = f->np_exact ? FV_OK : vpm_sync_count(f);
    if (rc) return rc;
    const int64_t before = f->np_host;
    const double *dbr = upload(ctx, S_BR, br, 3 * nr);
    const double *dfr = upload(ctx, S_FR, fr, 3 * nr);
    const double *dfl = upload(ctx, S_FL, fl, 3 * nr);
    const double *dbl = upload(ctx, S_BL, bl, 3 * nr);
    double *dg = upload(ctx, S_GAMMA, gamma, nr);
    if (!dbr || !dfr || !dfl || !dbl || !dg) return FV_E_HIP;
    int32_t zero = 0;
    FV_HIP_CHECK(ctx, hipMemcpyAsync(f->overflow_dev, &zero, sizeof zero, hipMemcpyHostToDevice,
                                     ctx->stream));
    rc = vpm_convert_rings(f, nr, dbr, dfr, dfl, dbl, dg, sigma);
    if (rc) return rc;
    int32_t ovf = 0;
    FV_HIP_CHECK(ctx, hipMemcpyAsync(&ovf, f->overflow_dev, sizeof ovf, hipMemcpyDeviceToHost,
                                     ctx->stream));
    FV_HIP_CHECK(ctx, hipMemcpyAsync(gamma, dg, sizeof(double) * nr, hipMemcpyDeviceToHost,
                                     ctx->stream));
    rc = vpm_sync_count(f);
    if (rc) return rc;
    if (ovf)
        return set_error(ctx, FV_E_CAP, "Particle field would overflow (capacity %lld)",
                         (long long)f->cap);
    if (added) *added = f->np_host - before;
    return FV_OK;
}

}  // extern "C"
