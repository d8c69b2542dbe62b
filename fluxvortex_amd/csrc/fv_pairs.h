// fv_pairs.h -- device math of the pair interactions and the tiled pair-sum
// kernel shared by every source/target sum in the library.
//
// Mapping (gfx950, wave64): TPT targets per lane (1, 2 or 4); a workgroup of
// BLOCK lanes stages BLOCK source records at a time in LDS from coalesced HBM
// reads (per-source constants formed once there) and every lane walks the
// tile in the same order (LDS broadcast reads).  Each lane keeps its own
// accumulators, so sums are atomic-free and bitwise reproducible.  Large
// source sets are split over blockIdx.y into fixed chunks whose partial sums
// are added in chunk order by k_reduce_partials.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

#include "fv_gauss_table.h"

namespace fv {

constexpr double kInv4Pi = 0.07957747154594767;   // warp_kernels.py:20
constexpr double kFourPi = 12.566370614359172;
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

// One vortex particle with per-source constants, built when a tile is staged:
//   Winckelmans:  k0 = 1/sigma^2, k1 = -1/(128 pi sigma^3), k2 = 1.5 k1,
//                 k3 = 3/(512 pi sigma^5), k4 = 2.5 k3, k5 = sigma^2
//                 (k1..k4 carry the 2^-5 / 2^-7 of the doubled inverse square
//                 root Z = 2 t, see part_factors_f64)
//   Gaussian-erf: k0 = 1/sigma, k1 = sigma, k2 = 1/sigma^2, k3 = k4 = k5 = 0
// Mixed-precision records (make_part_rec<1>) hold the primed strengths
// ax..az = -alpha / 4pi of part_factors_mixed, keep k0 and k2 for the
// Gaussian, and overlay k3..k5 with floats: Winckelmans f[3] = sigma^2,
// f[4] = 2.5 sigma^2, f[5] = 10.5 sigma^2.
struct alignas(16) PartRec {
    double x, y, z, ax, ay, az;
    double k0, k1, k2;
    union {
        struct {
            double k3, k4, k5;
        };
        float f[6];
    };
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

// 2/sqrt(a) for normal a > 0: the Newton step of rsqrt_refined without its
// halving, y (3 - a y^2), three operations after the estimate.  Callers fold
// the factor 2 into their constants.
__device__ __forceinline__ double rsqrt2_refined(double a)
{
    const double y = __builtin_amdgcn_rsq(a);
    return y * fma(-(a * y), y, 3.0);
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
// Gaussian-erf factors G1 = g / rho^3, G2 = g' / rho^4 - 3 g / rho^5 and
// G3 = g' / rho^2 - g / rho^3 for rho = r / sigma in [0, 10], from the
// piecewise polynomials of fv_gauss_table.h (fitted at 50 digits; about one
// unit in the last place of the terms' scale).  The closed form g = erf -
// sqrt(2/pi) rho e^(-rho^2/2) cancels as rho -> 0 (g ~ rho^3), and the erf and
// exp it needs are long, branchy library routines that a wavefront runs for
// every interval its lanes fall in; the table costs one row of LDS reads and
// a degree-12 Horner chain per factor.  Every kernel that evaluates the
// Gaussian factors stages the table in LDS first (gauss_stage(), before a
// barrier).  rho beyond the table or NaN selects the last piece (callers
// discard such pairs).
static __shared__ __attribute__((aligned(16))) double g_gauss_lds[kGaussPieces * 3 * kGaussRow];

__device__ __forceinline__ void gauss_stage()
{
    const double *src = &kGaussTab[0][0][0];
    for (int i = threadIdx.x; i < kGaussPieces * 3 * kGaussRow; i += blockDim.x)
        g_gauss_lds[i] = src[i];
}

__device__ __forceinline__ double gauss_poly(const double *c, double t)
{
    double v = c[kGaussDeg];
#pragma unroll
    for (int j = kGaussDeg - 1; j >= 0; --j) v = fma(v, t, c[j]);
    return v;
}

// Table factors FA and, when FB >= 0, FB (0: G1, 1: G2, 2: G3) at rho.
template <int FA, int FB>
__device__ __forceinline__ void gauss_factors(double rho, double &a, double &b)
{
    int k = rho < 10.0 ? (int)(rho * 2.0) : kGaussPieces - 1;
    k = k < 0 ? 0 : (k > kGaussPieces - 1 ? kGaussPieces - 1 : k);
    const double t = fma(4.0, rho, -(double)(2 * k + 1));
    const double *c = g_gauss_lds + k * (3 * kGaussRow);
    a = gauss_poly(c + FA * kGaussRow, t);
    b = FB >= 0 ? gauss_poly(c + (FB >= 0 ? FB : 0) * kGaussRow, t) : 0.0;
}

// Smoothing-kernel factors for one pair at squared distance r2:
//   CG = -(1/4pi) g(rho)/r^3                       (velocity factor)
//   CH = -(1/4pi) (g'(rho)/(sigma r) - 3 g/r^2)/r^3  (gradient factor)
// so that u += CG (dx x alpha) and
//   J_ij += CH (dx x alpha)_i dx_j + CG eps_ijl alpha_l     (kernel.py:163-188).
// Pairs with keep == false (the self pair and coincident particles) get
// CG = CH = 0 by selection, so non-finite intermediates never leak.
// Winckelmans (kernel.py:42-48): with q = rho^2 + 1 and t = q^(-1/2),
//   g/r^3 = (q + 1.5) t^5 / sigma^3,  (g'/(sigma r) - 3g/r^2)/r^3 = -(3q + 7.5) t^7 / sigma^5;
// evaluated with Z = 2t from rsqrt2_refined, Z^5 = 32 t^5 and Z^7 = 128 t^7
// folded into k1..k4: CG = (k1 q + k2) Z^5, CH = (k3 q + k4) Z^7.
// Gaussian-erf (kernel.py:26-31): g = erf(rho/sqrt2) - sqrt(2/pi) rho e^(-rho^2/2),
//   g' = sqrt(2/pi) rho^2 e^(-rho^2/2); for rho <= 10 CG = -G1 / (4 pi sigma^3) and
//   CH = -G2 / (4 pi sigma^5) from the table above.
template <int KERNEL, bool GRAD>
__device__ __forceinline__ void part_factors_f64(double r2, bool keep, const PartRec &s,
                                                 double &CG, double &CH)
{
    if (KERNEL == 1) {
        // excluded pairs: r^2 replaced by a value in [2^512, 2^512 (1 + 2^-20))
        // (only the high word is selected, one instruction), so that Z^5
        // underflows to exactly 0 while q and the factor polynomials stay
        // finite for any core sigma >= 1e-6: CG = CH = 0 exactly
        const double r2k = keep ? r2 : __hiloint2double(0x5FF00000, __double2loint(r2));
        const double q = fma(r2k, s.k0, 1.0);
        const double Z = rsqrt2_refined(q);
        const double Z2 = Z * Z;
        const double Z5 = Z2 * Z2 * Z;
        CG = fma(q, s.k1, s.k2) * Z5;
        if (GRAD) CH = fma(q, s.k3, s.k4) * (Z5 * Z2);
    } else {
        if (r2 * s.k2 > 100.0) {
            // rho > 10: erf(rho/sqrt2) == 1 and sqrt(2/pi) rho e^(-rho^2/2) < 2e-21 in
            // FP64, so g == 1 and g' == 0 to the last bit of the full formula:
            // CG = -ir^3 / 4pi, CH = 3 ir^5 / 4pi, with ir = Z / 2.
            const double Z = rsqrt2_refined(r2);
            const double Z2 = Z * Z;
            const double Z3 = Z2 * Z;
            CG = (-kInv4Pi / 8.0) * Z3;
            if (GRAD) CH = (3.0 * kInv4Pi / 32.0) * (Z3 * Z2);
        } else {
            const double ir = 0.5 * rsqrt2_refined(r2);
            const double rho = (r2 * ir) * s.k0;
            double G1, G2;
            gauss_factors<0, GRAD ? 1 : -1>(rho, G1, G2);
            const double is3 = s.k0 * s.k2;   // 1 / sigma^3
            CG = (-kInv4Pi * G1) * is3;
            if (GRAD) CH = (-kInv4Pi * G2) * (is3 * s.k2);
        }
        CG = keep ? CG : 0.0;
        if (GRAD) CH = keep ? CH : 0.0;
    }
}

// FP32 factors, primed: CG' = -4pi CG and CH' = -4pi CH, which pair with the
// primed strengths a' = -a / 4pi of the record (CG' a' = CG a).  The Gaussian
// pairs with rho < 0.25 (where g cancels in FP32) are evaluated in FP64
// instead, from the FP64 differences.
template <int KERNEL, bool GRAD>
__device__ __forceinline__ void part_factors_mixed(float r2, bool keep, double dxd, double dyd,
                                                   double dzd, const PartRec &s, float &CG,
                                                   float &CH)
{
    if (KERNEL == 1) {
        // The factors in a form with no 1/sigma^n (which overflows FP32 once a
        // stretched core shrinks towards its 1e-6 floor): with
        // u = (r^2 + sigma^2)^(-1/2),
        //   (rho^2+2.5) t^5 / sigma^3      = (r^2 + 2.5 sigma^2) u^2 u^3,
        //   (3 rho^2+10.5) t^7 / sigma^5   = (3 r^2 + 10.5 sigma^2) u^2 u^5,
        // every partial product bounded for any sigma.
        float u = __builtin_amdgcn_rsqf(r2 + s.f[3]);   // finite: sigma > 0
        u = keep ? u : 0.0f;
        const float u2 = u * u;
        const float u3 = u2 * u;
        CG = ((r2 + s.f[4]) * u2) * u3;
        if (GRAD) CH = -(fmaf(3.0f, r2, s.f[5]) * u2) * (u3 * u2);
    } else {
        const float ir = __builtin_amdgcn_rsqf(r2);
        if (r2 * (float)s.k2 > 100.0f) {   // g == 1, g' == 0 (see the FP64 path)
            const float ir3 = ir * ir * ir;
            CG = ir3;
            if (GRAD) CH = -3.0f * ir3 * (ir * ir);
        } else {
            const float r = r2 * ir;
            const float rho = r * (float)s.k0;
            if (rho < 0.25f) {
                const double r2d = dxd * dxd + dyd * dyd + dzd * dzd;
                double cg, ch = 0.0;
                part_factors_f64<0, GRAD>(r2d, true, s, cg, ch);
                CG = (float)(-kFourPi * cg);
                if (GRAD) CH = (float)(-kFourPi * ch);
            } else {
                const float ev = __expf(-0.5f * rho * rho);
                const float g = erff(rho * (float)kInvSqrt2) - (float)kSqrt2OverPi * rho * ev;
                const float ir3 = ir * ir * ir;
                CG = g * ir3;
                if (GRAD) {
                    const float dgdr = (float)kSqrt2OverPi * rho * rho * ev;
                    CH = (dgdr * ((float)s.k0 * ir) - 3.0f * g * (ir * ir)) * ir3;
                }
            }
        }
        CG = keep ? CG : 0.0f;
        if (GRAD) CH = keep ? CH : 0.0f;
    }
}

// Gaussian factors of the vorticity pair omega = CH (dx . a_s) dx - B a_s,
// B = CH r^2 + 2 CG, FP64.  For rho > 10 (g == 1, g' == 0 as in
// part_factors_f64) CH r^2 = 3/4pi ir^3 and 2 CG = -2/4pi ir^3, so
// B = ir^3 / 4pi = Z^3 / (32 pi): one product instead of the sum.
__device__ __forceinline__ void gauss_curl_factors_f64(double r2, bool keep, const PartRec &s,
                                                       double &CH, double &B)
{
    if (r2 * s.k2 > 100.0) {
        const double Z = rsqrt2_refined(r2);
        const double Z2 = Z * Z;
        const double Z3 = Z2 * Z;
        CH = (3.0 * kInv4Pi / 32.0) * (Z3 * Z2);
        B = (kInv4Pi / 8.0) * Z3;
    } else {
        // CH = -G2 / (4 pi sigma^5), B = -G3 / (4 pi sigma^3) (table)
        const double ir = 0.5 * rsqrt2_refined(r2);
        const double rho = (r2 * ir) * s.k0;
        double G2, G3;
        gauss_factors<1, 2>(rho, G2, G3);
        const double is3 = s.k0 * s.k2;
        CH = (-kInv4Pi * G2) * (is3 * s.k2);
        B = (-kInv4Pi * G3) * is3;
    }
    CH = keep ? CH : 0.0;
    B = keep ? B : 0.0;
}

// The reference CPU path's Gaussian factors at r = 1e-12 (kernel.py:156:
// r = sqrt(max(r^2, 1e-24))), used for the self pair and exactly coincident
// particles (pair rule 1): CG = -(1/4pi) g / r^3, CH = -(1/4pi) aux / r^3 with
// aux = g'/(sigma r) - 3 g / r^2.  There g = erf(rho/sqrt2) - sqrt(2/pi) rho
// e^(-rho^2/2) is pure round-off, and its sign decides whether the
// relaxation turns a particle's alpha against itself, so it is evaluated with
// the reference's operations in the reference's order and no contraction:
// erf as scipy.special.erf computes it below 1 (Cephes: x T(x^2) / U(x^2),
// Horner, the published coefficients), IEEE divisions, 1/r^3 = 1e36 and
// r^2 = 1e-24 exactly as NumPy forms them.  rho = 1e-12 / sigma < 1 for every
// core sigma >= 1e-6 the field allows.
__device__ inline double erf_cephes_below1(double x)
{
#pragma clang fp contract(off)
    const double z = x * x;
    double t = 9.60497373987051638749E0;
    t = t * z + 9.00260197203842689217E1;
    t = t * z + 2.23200534594684319226E3;
    t = t * z + 7.00332514112805075473E3;
    t = t * z + 5.55923013010394962768E4;
    double u = z + 3.35617141647503099647E1;
    u = u * z + 5.21357949780152679795E2;
    u = u * z + 4.59432382970980127987E3;
    u = u * z + 2.26290000613890934246E4;
    u = u * z + 4.92673942608635921086E4;
    return x * t / u;
}

__device__ inline void gauss_clamped_factors(double sigma, double &CG, double &CH)
{
#pragma clang fp contract(off)
    const double r = 1.0e-12;
    const double rb = r / sigma;
    const double rb2 = rb * rb;
    const double ev = exp(-0.5 * rb2);
    const double g = erf_cephes_below1(rb / 1.4142135623730951) - (kSqrt2OverPi * rb) * ev;
    const double dgdr = (kSqrt2OverPi * rb2) * ev;
    const double aux = dgdr / (sigma * r) - (3.0 * g) / 1.0e-24;
    CG = (-kInv4Pi * g) * 1.0e36;
    CH = (-kInv4Pi * aux) * 1.0e36;
}

template <int PREC>
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
        r.k1 = (-kInv4Pi / 32.0) * is3;
        r.k2 = 1.5 * r.k1;
        if (PREC == 0) {
            r.k3 = (3.0 * kInv4Pi / 128.0) * (is3 * is2);
            r.k4 = 2.5 * r.k3;
            r.k5 = 0.0;
        }
    } else {
        r.k0 = 1.0 / sigma;
        r.k1 = sigma;
        r.k2 = r.k0 * r.k0;
        if (PREC == 0) r.k3 = r.k4 = r.k5 = 0.0;
    }
    if (PREC == 1) {
        r.ax = -kInv4Pi * ax;
        r.ay = -kInv4Pi * ay;
        r.az = -kInv4Pi * az;
        const double sg2 = sigma * sigma;
        r.f[3] = (float)sg2;
        r.f[4] = (float)(2.5 * sg2);
        r.f[5] = (float)(10.5 * sg2);
    }
    return r;
}

// ------------------------------------------------------------ pair ops
// Each op defines: Args, Src, Tgt, Acc, kComp, load_src, load_tgt, init (zero
// accumulators), pair, finish (per-target epilogue before the sums are stored).

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
    __device__ static void init(Acc &c) { c.v[0] = c.v[1] = c.v[2] = 0.0; }
    __device__ static void finish(const Tgt &, Acc &) {}
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

template <int PREC>
__device__ __forceinline__ PartRec load_part(const PartSrcArgs &p, int64_t j)
{
    const int64_t k = j * p.stride;
    return make_part_rec<PREC>(p.kernel, p.x[k], p.y[k], p.z[k], p.ax[k], p.ay[k], p.az[k],
                               p.sig[j]);
}

struct PartTgtArgs {
    const double *x, *y, *z;       // target positions, SoA or AoS (stride 3)
    const double *ax, *ay, *az;    // target alphas (stretching only)
    int stride;
};

// mode: 0 velocity (3), 1 Jacobian (9), 2 stretching S = J alpha_t (3),
//       3 vorticity (curl, 3), 4 vorticity of the Gaussian field under the
//       reference CPU path's pair rule (kernel.py:156): pairs closer than
//       1e-12 are not skipped but take the factors of r = 1e-12
//       (gauss_clamped_factors, carried in the record's k3 / k4; primed by
//       -4pi in mixed records), omega += CH (dx . a) dx - (CH r^2 + 2 CG) a.
//       Those clamped terms (~1e8 |a|) are summed apart from the others
//       (components 3..5 of the sums, added by result() at the very end), so
//       the terms of coincident particles of opposite alpha cancel exactly in
//       any tile / chunk order instead of leaving round-off of their size.
template <int KERNEL, int PREC, int MODE>
struct PartOp {
    static_assert(MODE != 4 || KERNEL == 0, "mode 4 is the Gaussian vorticity");
    static constexpr int kComp = (MODE == 1) ? 9 : (MODE == 4) ? 6 : 3;
    static constexpr int kPrec = PREC;
    // source-loop unroll at two targets per lane: the Gaussian's erf / exp
    // branch doubled would exceed 128 VGPRs (three wavefronts per SIMD)
    static constexpr int kUnroll2 = (KERNEL == 0) ? 1 : 2;
    static constexpr bool kGaussTab = KERNEL == 0;   // gauss_stage() before the pairs
    struct Args {
        PartSrcArgs src;
        PartTgtArgs tgt;
    };
    using Src = PartRec;
    struct Tgt {
        double x, y, z, ax, ay, az;
    };
    // v: the sums; g: sum of CG alpha_s (stretching / Jacobian), whose cross
    // product with the target's alpha finish() adds once per target.
    struct Acc {
        double v[kComp];
        double g[3];
    };
    // a source record of this op (the clamped-pair factors for mode 4)
    __device__ static Src make(int kernel, double x, double y, double z, double ax, double ay,
                               double az, double sigma)
    {
        PartRec r = make_part_rec<PREC>(kernel, x, y, z, ax, ay, az, sigma);
        if (MODE == 4) {
            double cg, ch;
            gauss_clamped_factors(sigma, cg, ch);
            const double pf = PREC == 0 ? 1.0 : -kFourPi;
            r.k3 = pf * ch;
            r.k4 = pf * (cg + cg);
        }
        return r;
    }
    __device__ static Src load_src(const Args &a, int64_t j)
    {
        const int64_t k = j * a.src.stride;
        return make(a.src.kernel, a.src.x[k], a.src.y[k], a.src.z[k], a.src.ax[k], a.src.ay[k],
                    a.src.az[k], a.src.sig[j]);
    }
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
    __device__ static void init(Acc &c)
    {
#pragma unroll
        for (int k = 0; k < kComp; ++k) c.v[k] = 0.0;
        c.g[0] = c.g[1] = c.g[2] = 0.0;
    }
    // the k-th component (k < 3; mode 1: k < 9) of a finished sum
    __device__ static double result(const Acc &c, int k)
    {
        return MODE == 4 ? c.v[k] + c.v[k + 3] : c.v[k];
    }
    // S = sum CH (dx . a_t)(dx x a_s) + a_t x sum CG a_s  (stretching);
    // J_ij = sum CH (dx x a_s)_i dx_j + eps_ijl (sum CG a_s)_l  (Jacobian).
    __device__ static void finish(const Tgt &t, Acc &c)
    {
        if (MODE == 2) {
            c.v[0] += t.ay * c.g[2] - t.az * c.g[1];
            c.v[1] += t.az * c.g[0] - t.ax * c.g[2];
            c.v[2] += t.ax * c.g[1] - t.ay * c.g[0];
        } else if (MODE == 1) {
            c.v[1] += c.g[2]; c.v[2] -= c.g[1];
            c.v[3] -= c.g[2]; c.v[5] += c.g[0];
            c.v[6] += c.g[1]; c.v[7] -= c.g[0];
        }
    }
    __device__ static void pair(const Tgt &t, const Src &s, Acc &c)
    {
        const double dxd = t.x - s.x, dyd = t.y - s.y, dzd = t.z - s.z;
        if (PREC == 0) {
            const double r2d = dxd * dxd + dyd * dyd + dzd * dzd;
            const bool keep = r2d >= 1.0e-24;   // self pair excluded, warp_vpm.py:37-38
            double CG, CH = 0.0;
            if (KERNEL == 0 && (MODE == 3 || MODE == 4)) {
                double B;
                gauss_curl_factors_f64(r2d, keep, s, CH, B);
                if constexpr (MODE == 4) {
                    if (!keep) clamped_pair(dxd, dyd, dzd, r2d, s, c);
                }
                const double a = CH * (dxd * s.ax + dyd * s.ay + dzd * s.az);
                c.v[0] = fma(-B, s.ax, fma(a, dxd, c.v[0]));
                c.v[1] = fma(-B, s.ay, fma(a, dyd, c.v[1]));
                c.v[2] = fma(-B, s.az, fma(a, dzd, c.v[2]));
                return;
            }
            part_factors_f64<KERNEL, MODE != 0>(r2d, keep, s, CG, CH);
            accumulate(dxd, dyd, dzd, r2d, CG, CH, s, t, c);
        } else {
            // FP32 kernel factors (the rsqrt / erf / exp work), FP64 geometry and
            // sums: on gfx950 without packed FP32 an FP32 operation issues at the
            // FP64 rate, so FP32 pays only where it removes work (no Newton
            // step, FP32 erf / exp), and FP64 sums keep exactly cancelling pair
            // terms (coincident particles of opposite strength) cancelling.
            const double r2d = dxd * dxd + dyd * dyd + dzd * dzd;
            const bool keep = r2d >= 1.0e-24;
            float CGf, CHf = 0.0f;
            part_factors_mixed<KERNEL, MODE != 0>((float)r2d, keep, dxd, dyd, dzd, s, CGf, CHf);
            if constexpr (MODE == 4) {
                if (!keep) clamped_pair(dxd, dyd, dzd, r2d, s, c);
            }
            accumulate(dxd, dyd, dzd, r2d, (double)CGf, (double)CHf, s, t, c);
        }
    }
    // a pair closer than 1e-12 under the clamped rule (mode 4), into the
    // separate components 3..5; rare (the self pair, coincident particles)
    __device__ static void clamped_pair(double dx, double dy, double dz, double r2,
                                        const Src &s, Acc &c)
    {
        const double B = fma(s.k3, r2, s.k4);
        const double a = s.k3 * (dx * s.ax + dy * s.ay + dz * s.az);
        c.v[3] = fma(-B, s.ax, fma(a, dx, c.v[3]));
        c.v[4] = fma(-B, s.ay, fma(a, dy, c.v[4]));
        c.v[5] = fma(-B, s.az, fma(a, dz, c.v[5]));
    }
    __device__ static void accumulate(double dx, double dy, double dz, double r2, double CG,
                                      double CH, const Src &s, const Tgt &t, Acc &c)
    {
        if (MODE == 3 || MODE == 4) {
            // omega = CH dx x (dx x a_s) - 2 CG a_s = CH (dx . a_s) dx - (CH r^2 + 2 CG) a_s
            const double a = CH * (dx * s.ax + dy * s.ay + dz * s.az);
            const double b = fma(CH, r2, CG + CG);
            c.v[0] = fma(-b, s.ax, fma(a, dx, c.v[0]));
            c.v[1] = fma(-b, s.ay, fma(a, dy, c.v[1]));
            c.v[2] = fma(-b, s.az, fma(a, dz, c.v[2]));
            return;
        }
        const double cx = dy * s.az - dz * s.ay;
        const double cy = dz * s.ax - dx * s.az;
        const double cz = dx * s.ay - dy * s.ax;
        if (MODE == 0) {
            c.v[0] = fma(CG, cx, c.v[0]);
            c.v[1] = fma(CG, cy, c.v[1]);
            c.v[2] = fma(CG, cz, c.v[2]);
        } else if (MODE == 1) {
            const double hx = CH * cx, hy = CH * cy, hz = CH * cz;
            c.v[0] = fma(hx, dx, c.v[0]);
            c.v[1] = fma(hx, dy, c.v[1]);
            c.v[2] = fma(hx, dz, c.v[2]);
            c.v[3] = fma(hy, dx, c.v[3]);
            c.v[4] = fma(hy, dy, c.v[4]);
            c.v[5] = fma(hy, dz, c.v[5]);
            c.v[6] = fma(hz, dx, c.v[6]);
            c.v[7] = fma(hz, dy, c.v[7]);
            c.v[8] = fma(hz, dz, c.v[8]);
            c.g[0] = fma(CG, s.ax, c.g[0]);
            c.g[1] = fma(CG, s.ay, c.g[1]);
            c.g[2] = fma(CG, s.az, c.g[2]);
        } else {
            const double h = CH * (dx * t.ax + dy * t.ay + dz * t.az);
            c.v[0] = fma(h, cx, c.v[0]);
            c.v[1] = fma(h, cy, c.v[1]);
            c.v[2] = fma(h, cz, c.v[2]);
            c.g[0] = fma(CG, s.ax, c.g[0]);
            c.g[1] = fma(CG, s.ay, c.g[1]);
            c.g[2] = fma(CG, s.az, c.g[2]);
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
// Source-loop unroll at two targets per lane: the op's kUnroll2, else 2.
template <class Op, class = void>
struct PairUnroll2 {
    static constexpr int value = 2;
};

// Ops whose pairs evaluate the Gaussian factor table (staged in LDS).
template <class Op, class = void>
struct UsesGaussTab {
    static constexpr bool value = false;
};
template <class Op>
struct UsesGaussTab<Op, std::void_t<decltype(Op::kGaussTab)>> {
    static constexpr bool value = Op::kGaussTab;
};
template <class Op>
struct PairUnroll2<Op, decltype((void)Op::kUnroll2)> {
    static constexpr int value = Op::kUnroll2;
};

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
    // chunk <= 0: the slices share the device count evenly (whole tiles)
    if (chunk <= 0) chunk = ((n_src + gridDim.y - 1) / gridDim.y + BLOCK - 1) / BLOCK * BLOCK;
    if constexpr (UsesGaussTab<Op>::value) gauss_stage();   // read after the tile barrier
    typename Op::Tgt t[TPT];
    typename Op::Acc acc[TPT];
#pragma unroll
    for (int k = 0; k < TPT; ++k) {
        const int64_t i = i0 + threadIdx.x + k * BLOCK;
        t[k] = Op::load_tgt(a, i < n_tgt ? i : i0);
        Op::init(acc[k]);
    }
    const int64_t s0 = (int64_t)blockIdx.y * chunk;
    const int64_t s1 = (s0 + chunk < n_src) ? s0 + chunk : n_src;
    for (int64_t base = s0; base < s1; base += BLOCK) {
        const int cnt = (int)((s1 - base < BLOCK) ? (s1 - base) : BLOCK);
        __syncthreads();
        if (threadIdx.x < cnt) tile[threadIdx.x] = Op::load_src(a, base + threadIdx.x);
        __syncthreads();
#pragma unroll(TPT == 1 ? 2 : TPT == 2 ? PairUnroll2<Op>::value : 1)
        for (int j = 0; j < cnt; ++j) {
            const typename Op::Src src = tile[j];
#pragma unroll
            for (int k = 0; k < TPT; ++k) Op::pair(t[k], src, acc[k]);
        }
    }
#pragma unroll
    for (int k = 0; k < TPT; ++k) {
        const int64_t i = i0 + threadIdx.x + k * BLOCK;
        Op::finish(t[k], acc[k]);
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

// Many slices (a few targets over many sources): one wavefront per output
// element, lane l adding slices l, l + 64, ... in order, then a fixed
// butterfly over the lanes (deterministic; 4 outputs per 256-lane group).
template <int COMP>
__global__ __launch_bounds__(256) void k_reduce_partials_wide(const double *__restrict__ partial,
                                                              int64_t n_arg,
                                                              const int64_t *__restrict__ dyn_n,
                                                              int64_t stride, int split,
                                                              double *__restrict__ out)
{
    const int64_t n = dyn_n ? *dyn_n : n_arg;
    const int64_t e = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (e >= n * COMP) return;   // uniform over the wavefront
    const int lane = threadIdx.x & 63;
    const int64_t i = e / COMP, c = e % COMP;
    double s = 0.0;
    for (int y = lane; y < split; y += 64) s += partial[((int64_t)y * stride + i) * COMP + c];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if (lane == 0) out[i * COMP + c] = s;
}

}  // namespace fv
