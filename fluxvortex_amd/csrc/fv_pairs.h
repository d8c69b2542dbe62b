// fv_pairs.h -- device math of the pair interactions and the tiled pair-sum
// kernel shared by every source/target sum in the library.
//
// Mapping (gfx950, wave64): one target per lane; a workgroup of BLOCK lanes
// stages BLOCK source records at a time in LDS from coalesced HBM reads and
// every lane walks the tile in the same order (LDS broadcast reads).  Each lane
// keeps its own accumulators, so sums are atomic-free and bitwise
// reproducible.  Large source sets are split over blockIdx.y into fixed chunks
// whose partial sums are added in chunk order by k_reduce_partials.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace fv {

constexpr double kInv4Pi = 0.07957747154594767;   // warp_kernels.py:20
constexpr double kFourLamb = 5.02572;             // warp_kernels.py:18 (4 * 1.25643)
constexpr double kSquire = 1.0e-4;                // warp_kernels.py:19
constexpr double kSqrt2OverPi = 0.7978845608028654;   // warp_vpm.py:14
constexpr double kInvSqrt2 = 0.7071067811865476;      // warp_vpm.py:13

// ----------------------------------------------------------------- records
// One straight vortex segment, preprocessed once per call:
//   c1 = strength / 4pi, c2 = |r0|^2 rc^2, tol = 1e-10 |r0|
// (tol = +inf for |r0| < 2.22e-16 so that the segment is always skipped,
// warp_kernels.py:52-53).
struct alignas(16) SegRec {
    double sx, sy, sz, ex, ey, ez;
    double c1, c2, tol, pad;
};

// One vortex particle with per-source constants:
//   Winckelmans:  k0 = 1/sigma^2, k1 = -1/(4 pi sigma^3), k2 = 1/(4 pi sigma^5)
//   Gaussian-erf: k0 = 1/sigma,   k1 = sigma,             k2 = 1/sigma^2
struct alignas(16) PartRec {
    double x, y, z, ax, ay, az;
    double k0, k1, k2, pad;
};

struct Pt3 {
    double x, y, z;
};

// ------------------------------------------------------------ FP64 helpers
// 1/sqrt(a) for normal a > 0: the v_rsq_f64 estimate (relative error <= 5.3e-8)
// refined by one Newton step, which squares the error: measured max relative
// error 4.1e-15 over a in [1e-20, 1e20] (2^20 samples).  No denormal scaling,
// so a must be a normal number.
__device__ __forceinline__ double rsqrt_refined(double a)
{
    const double y = __builtin_amdgcn_rsq(a);
    const double r = fma(-a * y, y, 1.0);
    return fma(y, 0.5 * r, y);
}

// 1/d for normal d: the v_rcp_f64 estimate refined by one Newton step
// (measured max relative error 2.2e-15 over the same range).
__device__ __forceinline__ double rcp_refined(double d)
{
    const double y = __builtin_amdgcn_rcp(d);
    const double e = fma(-d, y, 1.0);
    return fma(y, e, y);
}

// ------------------------------------------------------- segment, FP64
// Ptera-form segment velocity (warp_kernels.py:46-96), rewritten for gfx950:
//   c4 = c1 (|r1| + |r2|)(|r1||r2| - r1.r2) / (|r1||r2| (|r1 x r2|^2 + c2))
//      = c1 (1/|r1| + 1/|r2|)(|r1||r2| - r1.r2) / (|r1 x r2|^2 + c2)
// with 1/|r1|, 1/|r2| from v_rsq_f64 + Newton and one refined reciprocal, so
// there is no IEEE sqrt/div expansion in the pair loop.  The skip rules are the
// reference's (:72-82) evaluated on squares (|r1|^2 < tol^2, |r2|^2 < tol^2,
// |r1 x r2|^2 < 1e-20 |r1|^2 |r2|^2); they differ from the reference's sqrt
// tests only when both sides round to the same value.  Skipped pairs are
// selected away, so their (possibly non-finite) intermediate values never
// reach the accumulator.
__device__ __forceinline__ void seg_pair_f64(double px, double py, double pz, const SegRec &s,
                                             double &ax, double &ay, double &az)
{
    const double r1x = s.sx - px, r1y = s.sy - py, r1z = s.sz - pz;
    const double r2x = s.ex - px, r2y = s.ey - py, r2z = s.ez - pz;
    const double r3x = r1y * r2z - r1z * r2y;
    const double r3y = r1z * r2x - r1x * r2z;
    const double r3z = r1x * r2y - r1y * r2x;
    const double r1sq = r1x * r1x + r1y * r1y + r1z * r1z;
    const double r2sq = r2x * r2x + r2y * r2y + r2z * r2z;
    const double r3sq = r3x * r3x + r3y * r3y + r3z * r3z;
    const double c3 = r1x * r2x + r1y * r2y + r1z * r2z;
    const double tol2 = s.tol * s.tol;
    const bool skip = (r1sq < tol2) || (r2sq < tol2) || (r3sq < 1.0e-20 * (r1sq * r2sq));
    const double i1 = rsqrt_refined(r1sq);
    const double i2 = rsqrt_refined(r2sq);
    const double r1r2 = (r1sq * i1) * (r2sq * i2);
    double c4 = s.c1 * (i1 + i2) * (r1r2 - c3) * rcp_refined(r3sq + s.c2);
    c4 = skip ? 0.0 : c4;
    ax = fma(c4, r3x, ax);
    ay = fma(c4, r3y, ay);
    az = fma(c4, r3z, az);
}

// ------------------------------------------------------- segment, mixed
// FP32 pair term with FP64 differences and FP64 accumulation.  Uses the
// cancellation-free numerator |r1||r2| - r1.r2 = |r1 x r2|^2 / (|r1||r2| + r1.r2)
// when r1.r2 > 0 (far field), and falls back to the FP64 pair when the target
// lies within 1e-2 rad of the segment's line (|r1 x r2| loses digits in FP32
// there) or closer than 1e-3 segment lengths to an end point.
__device__ __forceinline__ void seg_pair_mixed(double px, double py, double pz, const SegRec &s,
                                               double &ax, double &ay, double &az)
{
    const float r1x = (float)(s.sx - px), r1y = (float)(s.sy - py), r1z = (float)(s.sz - pz);
    const float r2x = (float)(s.ex - px), r2y = (float)(s.ey - py), r2z = (float)(s.ez - pz);
    const float r3x = r1y * r2z - r1z * r2y;
    const float r3y = r1z * r2x - r1x * r2z;
    const float r3z = r1x * r2y - r1y * r2x;
    const float r1s = r1x * r1x + r1y * r1y + r1z * r1z;
    const float r2s = r2x * r2x + r2y * r2y + r2z * r2z;
    const float r1 = __builtin_amdgcn_sqrtf(r1s);
    const float r2 = __builtin_amdgcn_sqrtf(r2s);
    const float r3sq = r3x * r3x + r3y * r3y + r3z * r3z;
    const float r1r2 = r1 * r2;
    const float tolf = (float)(s.tol * 1.0e7);   // 1e-3 |r0|
    const bool near = (r3sq < 1.0e-4f * r1r2 * r1r2) || (r1 < tolf) || (r2 < tolf);
    if (near) {
        seg_pair_f64(px, py, pz, s, ax, ay, az);
        return;
    }
    const float c3 = r1x * r2x + r1y * r2y + r1z * r2z;
    const bool far = c3 > 0.0f;
    const float num = far ? r3sq : (r1r2 - c3);
    const float dext = far ? (r1r2 + c3) : 1.0f;
    const float den = dext * r1r2 * (r3sq + (float)s.c2);
    const float c4 = (float)s.c1 * (r1 + r2) * num * __builtin_amdgcn_rcpf(den);
    ax += (double)(c4 * r3x);
    ay += (double)(c4 * r3y);
    az += (double)(c4 * r3z);
}

// ----------------------------------------------------- particle kernels
// Smoothing-kernel factors for one pair at squared distance r2 (> 0):
//   CG = -(1/4pi) g(rho)/r^3                       (velocity factor)
//   CH = -(1/4pi) (g'(rho)/(sigma r) - 3 g/r^2)/r^3  (gradient factor)
// so that u += CG (dx x alpha) and
//   J_ij += CH (dx x alpha)_i dx_j + CG eps_ijl alpha_l     (kernel.py:163-188).
// Winckelmans (kernel.py:42-48): with t = (rho^2+1)^(-1/2),
//   g/r^3 = (rho^2+2.5) t^5 / sigma^3,  (g'/(sigma r) - 3g/r^2)/r^3 = -(3 rho^2 + 10.5) t^7 / sigma^5.
// Gaussian-erf (kernel.py:26-31): g = erf(rho/sqrt2) - sqrt(2/pi) rho e^(-rho^2/2),
//   g' = sqrt(2/pi) rho^2 e^(-rho^2/2).
template <int KERNEL, bool GRAD>
__device__ __forceinline__ void part_factors_f64(double r2, const PartRec &s, double &CG,
                                                 double &CH)
{
    if (KERNEL == 1) {
        const double rho2 = r2 * s.k0;
        const double t = rsqrt_refined(rho2 + 1.0);
        const double t2 = t * t;
        const double t5 = t2 * t2 * t;
        CG = s.k1 * (rho2 + 2.5) * t5;
        if (GRAD) CH = s.k2 * fma(3.0, rho2, 10.5) * (t5 * t2);
    } else if (r2 * s.k2 > 100.0) {
        // rho > 10: erf(rho/sqrt2) == 1 and sqrt(2/pi) rho e^(-rho^2/2) < 2e-21 in
        // FP64, so g == 1 and g' == 0 to the last bit of the full formula.
        const double ir = rsqrt_refined(r2);
        const double ir2 = ir * ir;
        const double ir3 = ir2 * ir;
        CG = -kInv4Pi * ir3;
        if (GRAD) CH = 3.0 * kInv4Pi * ir3 * ir2;
    } else {
        const double r = sqrt(r2);
        const double rho = r * s.k0;
        const double ev = exp(-0.5 * rho * rho);
        const double g = erf(rho * kInvSqrt2) - kSqrt2OverPi * rho * ev;
        const double ir = 1.0 / r;
        const double ir3 = ir * ir * ir;
        CG = -kInv4Pi * g * ir3;
        if (GRAD) {
            const double dgdr = kSqrt2OverPi * rho * rho * ev;
            const double aux = dgdr * (s.k0 * ir) - 3.0 * g * (ir * ir);
            CH = -kInv4Pi * aux * ir3;
        }
    }
}

// FP32 factors; the Gaussian pairs with rho < 0.25 (where g cancels in FP32)
// are evaluated in FP64 instead.
template <int KERNEL, bool GRAD>
__device__ __forceinline__ void part_factors_mixed(float r2, double r2d, const PartRec &s,
                                                   float &CG, float &CH)
{
    if (KERNEL == 1) {
        // The same factors in a form with no 1/sigma^n (which overflows FP32
        // once a stretched core shrinks towards its 1e-6 floor): with
        // u = (r^2 + sigma^2)^(-1/2),
        //   (rho^2+2.5) t^5 / sigma^3      = (r^2 + 2.5 sigma^2) u^2 u^3,
        //   (3 rho^2+10.5) t^7 / sigma^5   = (3 r^2 + 10.5 sigma^2) u^2 u^5,
        // every factor bounded for any sigma.
        const float sg2 = (float)s.pad;
        const float u = __builtin_amdgcn_rsqf(r2 + sg2);
        const float u2 = u * u;
        const float u3 = u2 * u;
        CG = -(float)kInv4Pi * (fmaf(2.5f, sg2, r2) * u2) * u3;
        if (GRAD) CH = (float)kInv4Pi * (fmaf(10.5f, sg2, 3.0f * r2) * u2) * (u3 * u2);
    } else {
        const float ir = __builtin_amdgcn_rsqf(r2);
        if (r2 * (float)s.k2 > 100.0f) {   // g == 1, g' == 0 (see the FP64 path)
            const float ir3 = ir * ir * ir;
            CG = -(float)kInv4Pi * ir3;
            if (GRAD) CH = 3.0f * (float)kInv4Pi * ir3 * (ir * ir);
            return;
        }
        const float r = r2 * ir;
        const float rho = r * (float)s.k0;
        if (rho < 0.25f) {
            double cg, ch = 0.0;
            part_factors_f64<0, GRAD>(r2d, s, cg, ch);
            CG = (float)cg;
            if (GRAD) CH = (float)ch;
            return;
        }
        const float ev = __expf(-0.5f * rho * rho);
        const float g = erff(rho * (float)kInvSqrt2) - (float)kSqrt2OverPi * rho * ev;
        const float ir3 = ir * ir * ir;
        CG = -(float)kInv4Pi * g * ir3;
        if (GRAD) {
            const float dgdr = (float)kSqrt2OverPi * rho * rho * ev;
            const float aux = dgdr * ((float)s.k0 * ir) - 3.0f * g * (ir * ir);
            CH = -(float)kInv4Pi * aux * ir3;
        }
    }
}

__device__ __forceinline__ PartRec make_part_rec(int kernel, double x, double y, double z,
                                                 double ax, double ay, double az, double sigma)
{
    PartRec r;
    r.x = x; r.y = y; r.z = z;
    r.ax = ax; r.ay = ay; r.az = az;
    if (kernel == 1) {
        const double is = 1.0 / sigma;
        const double is2 = is * is;
        const double is3 = is2 * is;
        r.k0 = is2;
        r.k1 = -kInv4Pi * is3;
        r.k2 = kInv4Pi * is3 * is2;
        r.pad = sigma * sigma;   // the mixed-precision factors' form
    } else {
        r.k0 = 1.0 / sigma;
        r.k1 = sigma;
        r.k2 = r.k0 * r.k0;
        r.pad = 0.0;
    }
    return r;
}

// ------------------------------------------------------------ pair ops
// Each op defines: Args, Src, Tgt, kComp, load_src, load_tgt, pair, store.

// Segment sums at points (collapsed Biot-Savart).
template <int PREC>
struct SegOp {
    static constexpr int kComp = 3;
    static constexpr int kPrec = PREC;
    struct Args {
        const double *pts;    // AoS (n, 3)
        const SegRec *segs;
    };
    using Src = SegRec;
    using Tgt = Pt3;
    struct Acc {
        double v[3];
    };
    __device__ static Src load_src(const Args &a, int64_t j) { return a.segs[j]; }
    __device__ static Tgt load_tgt(const Args &a, int64_t i)
    {
        return Pt3{a.pts[3 * i], a.pts[3 * i + 1], a.pts[3 * i + 2]};
    }
    __device__ static void pair(const Tgt &t, const Src &s, Acc &c)
    {
        if (PREC == 0)
            seg_pair_f64(t.x, t.y, t.z, s, c.v[0], c.v[1], c.v[2]);
        else
            seg_pair_mixed(t.x, t.y, t.z, s, c.v[0], c.v[1], c.v[2]);
    }
};

// Particle sources stored SoA (device particle fields) or AoS (API calls).
struct PartSrcArgs {
    const double *x, *y, *z, *ax, *ay, *az, *sig;  // SoA, or AoS when stride == 3
    int stride;                                    // 1 (SoA) or 3 (AoS pos/alpha)
    int kernel;
};

__device__ __forceinline__ PartRec load_part(const PartSrcArgs &p, int64_t j)
{
    const int64_t k = j * p.stride;
    return make_part_rec(p.kernel, p.x[k], p.y[k], p.z[k], p.ax[k], p.ay[k], p.az[k], p.sig[j]);
}

struct PartTgtArgs {
    const double *x, *y, *z;       // target positions, SoA or AoS (stride 3)
    const double *ax, *ay, *az;    // target alphas (stretching only)
    int stride;
};

// mode: 0 velocity (3), 1 Jacobian (9), 2 stretching S = J alpha_t (3),
//       3 vorticity (curl, 3)
template <int KERNEL, int PREC, int MODE>
struct PartOp {
    static constexpr int kComp = (MODE == 1) ? 9 : 3;
    static constexpr int kPrec = PREC;
    struct Args {
        PartSrcArgs src;
        PartTgtArgs tgt;
    };
    using Src = PartRec;
    struct Tgt {
        double x, y, z, ax, ay, az;
    };
    struct Acc {
        double v[kComp];
    };
    __device__ static Src load_src(const Args &a, int64_t j) { return load_part(a.src, j); }
    __device__ static Tgt load_tgt(const Args &a, int64_t i)
    {
        const int64_t k = i * a.tgt.stride;
        Tgt t;
        t.x = a.tgt.x[k]; t.y = a.tgt.y[k]; t.z = a.tgt.z[k];
        if (MODE == 2) {
            t.ax = a.tgt.ax[k]; t.ay = a.tgt.ay[k]; t.az = a.tgt.az[k];
        } else {
            t.ax = t.ay = t.az = 0.0;
        }
        return t;
    }
    __device__ static void pair(const Tgt &t, const Src &s, Acc &c)
    {
        const double dxd = t.x - s.x, dyd = t.y - s.y, dzd = t.z - s.z;
        const double r2d = dxd * dxd + dyd * dyd + dzd * dzd;
        const bool keep = r2d >= 1.0e-24;   // self pair excluded, warp_vpm.py:37-38
        if (PREC == 0) {
            double CG, CH = 0.0;
            part_factors_f64<KERNEL, MODE != 0>(r2d, s, CG, CH);
            CG = keep ? CG : 0.0;
            CH = keep ? CH : 0.0;
            const double cx = dyd * s.az - dzd * s.ay;
            const double cy = dzd * s.ax - dxd * s.az;
            const double cz = dxd * s.ay - dyd * s.ax;
            accumulate(dxd, dyd, dzd, cx, cy, cz, CG, CH, s.ax, s.ay, s.az, t, c);
        } else {
            const float dx = (float)dxd, dy = (float)dyd, dz = (float)dzd;
            const float r2 = dx * dx + dy * dy + dz * dz;
            float CG, CH = 0.0f;
            part_factors_mixed<KERNEL, MODE != 0>(r2, r2d, s, CG, CH);
            CG = keep ? CG : 0.0f;
            CH = keep ? CH : 0.0f;
            const float fax = (float)s.ax, fay = (float)s.ay, faz = (float)s.az;
            const float cx = dy * faz - dz * fay;
            const float cy = dz * fax - dx * faz;
            const float cz = dx * fay - dy * fax;
            accumulate_mixed(dx, dy, dz, cx, cy, cz, CG, CH, fax, fay, faz, t, c);
        }
    }
    __device__ static void accumulate(double dx, double dy, double dz, double cx, double cy,
                                      double cz, double CG, double CH, double sax, double say,
                                      double saz, const Tgt &t, Acc &c)
    {
        if (MODE == 0) {
            c.v[0] = fma(CG, cx, c.v[0]);
            c.v[1] = fma(CG, cy, c.v[1]);
            c.v[2] = fma(CG, cz, c.v[2]);
        } else if (MODE == 1) {
            const double hx = CH * cx, hy = CH * cy, hz = CH * cz;
            const double gx = CG * sax, gy = CG * say, gz = CG * saz;
            c.v[0] = fma(hx, dx, c.v[0]);
            c.v[1] = fma(hx, dy, c.v[1]) + gz;
            c.v[2] = fma(hx, dz, c.v[2]) - gy;
            c.v[3] = fma(hy, dx, c.v[3]) - gz;
            c.v[4] = fma(hy, dy, c.v[4]);
            c.v[5] = fma(hy, dz, c.v[5]) + gx;
            c.v[6] = fma(hz, dx, c.v[6]) + gy;
            c.v[7] = fma(hz, dy, c.v[7]) - gx;
            c.v[8] = fma(hz, dz, c.v[8]);
        } else if (MODE == 2) {
            // S_i = CH c_i (dx . a_t) + CG (a_t x a_s)_i
            const double d = dx * t.ax + dy * t.ay + dz * t.az;
            const double h = CH * d;
            c.v[0] = fma(h, cx, c.v[0]) + CG * (t.ay * saz - t.az * say);
            c.v[1] = fma(h, cy, c.v[1]) + CG * (t.az * sax - t.ax * saz);
            c.v[2] = fma(h, cz, c.v[2]) + CG * (t.ax * say - t.ay * sax);
        } else {
            // omega = CH (dx x c) - 2 CG a_s
            c.v[0] = fma(CH, dy * cz - dz * cy, c.v[0]) - 2.0 * CG * sax;
            c.v[1] = fma(CH, dz * cx - dx * cz, c.v[1]) - 2.0 * CG * say;
            c.v[2] = fma(CH, dx * cy - dy * cx, c.v[2]) - 2.0 * CG * saz;
        }
    }
    __device__ static void accumulate_mixed(float dx, float dy, float dz, float cx, float cy,
                                            float cz, float CG, float CH, float sax, float say,
                                            float saz, const Tgt &t, Acc &c)
    {
        if (MODE == 0) {
            c.v[0] += (double)(CG * cx);
            c.v[1] += (double)(CG * cy);
            c.v[2] += (double)(CG * cz);
        } else if (MODE == 1) {
            const float hx = CH * cx, hy = CH * cy, hz = CH * cz;
            const float gx = CG * sax, gy = CG * say, gz = CG * saz;
            c.v[0] += (double)(hx * dx);
            c.v[1] += (double)(fmaf(hx, dy, gz));
            c.v[2] += (double)(fmaf(hx, dz, -gy));
            c.v[3] += (double)(fmaf(hy, dx, -gz));
            c.v[4] += (double)(hy * dy);
            c.v[5] += (double)(fmaf(hy, dz, gx));
            c.v[6] += (double)(fmaf(hz, dx, gy));
            c.v[7] += (double)(fmaf(hz, dy, -gx));
            c.v[8] += (double)(hz * dz);
        } else if (MODE == 2) {
            const float tax = (float)t.ax, tay = (float)t.ay, taz = (float)t.az;
            const float h = CH * (dx * tax + dy * tay + dz * taz);
            c.v[0] += (double)fmaf(h, cx, CG * (tay * saz - taz * say));
            c.v[1] += (double)fmaf(h, cy, CG * (taz * sax - tax * saz));
            c.v[2] += (double)fmaf(h, cz, CG * (tax * say - tay * sax));
        } else {
            c.v[0] += (double)fmaf(CH, dy * cz - dz * cy, -2.0f * CG * sax);
            c.v[1] += (double)fmaf(CH, dz * cx - dx * cz, -2.0f * CG * say);
            c.v[2] += (double)fmaf(CH, dx * cy - dy * cx, -2.0f * CG * saz);
        }
    }
};

// --------------------------------------------------- tiled pair-sum kernel
// Each lane owns TPT targets (i0 + lane + k * BLOCK); the workgroup covers
// BLOCK * TPT targets and walks its source chunk in LDS tiles of BLOCK
// records, so every staged record feeds TPT pairs per lane.
// partial[(blockIdx.y * n_tgt_stride + i) * kComp + c] = sum over this
// block's source chunk.  n_tgt / n_src come from *dyn_tgt / *dyn_src when
// those are non-null (device-side particle counts), else from the arguments.
template <class Op, int BLOCK, int TPT>
__global__ __launch_bounds__(BLOCK) void k_pairs(typename Op::Args a, int64_t n_tgt_arg,
                                                 int64_t n_src_arg,
                                                 const int64_t *__restrict__ dyn_tgt,
                                                 const int64_t *__restrict__ dyn_src,
                                                 int64_t chunk, int64_t n_tgt_stride,
                                                 double *__restrict__ partial)
{
    __shared__ typename Op::Src tile[BLOCK];
    const int64_t n_tgt = dyn_tgt ? *dyn_tgt : n_tgt_arg;
    const int64_t n_src = dyn_src ? *dyn_src : n_src_arg;
    const int64_t i0 = (int64_t)blockIdx.x * (BLOCK * TPT);
    if (i0 >= n_tgt) return;   // uniform over the block
    typename Op::Tgt t[TPT];
    typename Op::Acc acc[TPT];
#pragma unroll
    for (int k = 0; k < TPT; ++k) {
        const int64_t i = i0 + threadIdx.x + k * BLOCK;
        t[k] = Op::load_tgt(a, i < n_tgt ? i : i0);
#pragma unroll
        for (int c = 0; c < Op::kComp; ++c) acc[k].v[c] = 0.0;
    }
    const int64_t s0 = (int64_t)blockIdx.y * chunk;
    const int64_t s1 = (s0 + chunk < n_src) ? s0 + chunk : n_src;
    for (int64_t base = s0; base < s1; base += BLOCK) {
        const int cnt = (int)((s1 - base < BLOCK) ? (s1 - base) : BLOCK);
        __syncthreads();
        if (threadIdx.x < cnt) tile[threadIdx.x] = Op::load_src(a, base + threadIdx.x);
        __syncthreads();
#pragma unroll(TPT == 1 ? 2 : 1)
        for (int j = 0; j < cnt; ++j) {
            const typename Op::Src src = tile[j];
#pragma unroll
            for (int k = 0; k < TPT; ++k) Op::pair(t[k], src, acc[k]);
        }
    }
#pragma unroll
    for (int k = 0; k < TPT; ++k) {
        const int64_t i = i0 + threadIdx.x + k * BLOCK;
        if (i < n_tgt) {
            double *o = partial + ((int64_t)blockIdx.y * n_tgt_stride + i) * Op::kComp;
#pragma unroll
            for (int c = 0; c < Op::kComp; ++c) o[c] = acc[k].v[c];
        }
    }
}

// out[i][c] = sum_{y < split} partial[y][i][c], in y order.
template <int COMP>
__global__ void k_reduce_partials(const double *__restrict__ partial, int64_t n_arg,
                                  const int64_t *__restrict__ dyn_n, int64_t stride, int split,
                                  double *__restrict__ out)
{
    const int64_t n = dyn_n ? *dyn_n : n_arg;
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n * COMP) return;
    const int64_t i = e / COMP, c = e % COMP;
    double s = 0.0;
    for (int y = 0; y < split; ++y) s += partial[((int64_t)y * stride + i) * COMP + c];
    out[i * COMP + c] = s;
}

}  // namespace fv
