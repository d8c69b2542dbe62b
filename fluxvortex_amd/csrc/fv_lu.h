// fv_lu.h -- dense LU building blocks executed by one workgroup (device code
// shared by the batched circulation solve and the batched flutter sweep).
#pragma once

#include "fv_launch.h"

namespace fv {

// ---------------------------------------------------------------- blocked
// Right-looking blocked LU (panel width kNb) executed by one workgroup:
//   1. the n-k0 x kb panel is factored by wave 0 with partial pivoting
//      (wave-synchronous: no workgroup barrier inside the panel), swapping
//      the panel's rows and eliminating the right-hand side as it goes; when
//      the matrix lives in global memory the panel is first staged in LDS;
//   2. the panel's row swaps are applied to every other column (in order,
//      one column per thread), so L and U end in LAPACK's row order;
//   3. all waves form U12 = L11^{-1} A12 (one column per lane);
//   4. all waves update A22 -= L21 U12.
// The right-hand side and the pivots live in LDS; LDS row strides are odd so
// that the lanes of a column walk (one row per lane) hit distinct banks.
constexpr int kNb = 16;
constexpr int kLdp = kNb + 1;   // LDS row stride of a staged panel

// Factor reuse across solves of one system slot (all arrays per batch entry,
// device memory owned by the caller): the matrix the slot last factored
// (key, n x n FP64), its factors (fac, n x n of the factor type), the
// composed row order of its pivots ((P v)[i] = v[order[i]]) and a state word
// (1: key / fac / order describe a nonsingular factorization).  A solve whose
// matrix equals key bit for bit skips the factorization and takes fac and
// order instead; either way the solution comes from the same two
// substitutions (L y = P b, U x = y) on the same factors, so it is the same
// bit for bit.  key == nullptr: no reuse.
struct LuReuse {
    double *key;
    void *fac;
    int32_t *order;
    int32_t *state;
    int32_t *hits;   // nullable: solves that took the stored factors
};

// Bitwise inequality of two doubles (NaN payloads and signed zeros count).
__device__ __forceinline__ bool bits_differ(double x, double y)
{
    return __double_as_longlong(x) != __double_as_longlong(y);
}

// dst[r ldd + c] = src[r lds + c] over an n x n matrix by a workgroup of NT
// lanes, kU elements per lane in flight (32-bit indices: n <= 1024); dst
// nullable (compare only).  Returns whether any moved element differs bit for
// bit from cmp (dense, row stride n; nullable: false).
template <int NT, int kU = 8, typename TS, typename TD>
__device__ __forceinline__ bool mat_move(const TS *__restrict__ src, int lds, TD *__restrict__ dst,
                                         int ldd, int n, const double *__restrict__ cmp)
{
    const int nn = n * n;
    bool diff = false;
    for (int e0 = threadIdx.x; e0 < nn; e0 += kU * NT) {
        TS v[kU];
        double k[kU];
        int rr[kU], cc[kU];
#pragma unroll
        for (int j = 0; j < kU; ++j) {
            const int e = e0 + j * NT;
            rr[j] = e / n;
            cc[j] = e - rr[j] * n;
            v[j] = e < nn ? src[rr[j] * lds + cc[j]] : (TS)0;
            k[j] = (cmp && e < nn) ? cmp[e] : 0.0;
        }
#pragma unroll
        for (int j = 0; j < kU; ++j) {
            if (e0 + j * NT < nn) {
                if (dst) dst[rr[j] * ldd + cc[j]] = (TD)v[j];
                if (cmp) diff |= bits_differ((double)v[j], k[j]);
            }
        }
    }
    return diff;
}

// Final position of row i under the pivot interchanges perm[0..n) applied in
// order (row k swapped with row perm[k]); order[pivot_track(perm, n, i)] = i
// composes the sequence into one gather order, one row per lane.
__device__ __forceinline__ int pivot_track(const int *perm, int n, int i)
{
    int p = i;
    for (int k = 0; k < n; ++k) {
        const int q = perm[k];
        p = p == k ? q : (p == q ? k : p);
    }
    return p;
}

__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Wavefront reductions on DPP (row_shr 1/2/4/8 inside each 16-lane row, then
// row_bcast 15 / 31 across rows); lane 63 ends with the result, which
// v_readlane makes uniform.  Lanes whose DPP source is outside the row take
// the identity.
template <int CTRL, int ROWMASK>
__device__ __forceinline__ int dpp_i(int ident, int v)
{
    return __builtin_amdgcn_update_dpp(ident, v, CTRL, ROWMASK, 0xf, false);
}

template <int CTRL, int ROWMASK>
__device__ __forceinline__ double dpp_d(double ident, double v)
{
    const int lo = dpp_i<CTRL, ROWMASK>(__double2loint(ident), __double2loint(v));
    const int hi = dpp_i<CTRL, ROWMASK>(__double2hiint(ident), __double2hiint(v));
    return __hiloint2double(hi, lo);
}

__device__ __forceinline__ double readlane_d(double v, int lane)
{
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
    return __hiloint2double(hi, lo);
}

// Index of the largest |value| over the wavefront (ties: the smallest index);
// lanes without a candidate pass best = -1.  LAPACK's idamax rule.
__device__ __forceinline__ int wave_argmax(double best, int idx)
{
    double m = best;
    m = fmax(m, dpp_d<0x111, 0xf>(-1.0, m));
    m = fmax(m, dpp_d<0x112, 0xf>(-1.0, m));
    m = fmax(m, dpp_d<0x114, 0xf>(-1.0, m));
    m = fmax(m, dpp_d<0x118, 0xf>(-1.0, m));
    m = fmax(m, dpp_d<0x142, 0xa>(-1.0, m));
    m = fmax(m, dpp_d<0x143, 0xc>(-1.0, m));
    const double M = readlane_d(m, 63);
    const int big = 0x7fffffff;
    int c = (best == M) ? idx : big;
    c = min(c, dpp_i<0x111, 0xf>(big, c));
    c = min(c, dpp_i<0x112, 0xf>(big, c));
    c = min(c, dpp_i<0x114, 0xf>(big, c));
    c = min(c, dpp_i<0x118, 0xf>(big, c));
    c = min(c, dpp_i<0x142, 0xa>(big, c));
    c = min(c, dpp_i<0x143, 0xc>(big, c));
    return __builtin_amdgcn_readlane(c, 63);
}

__device__ __forceinline__ double readlane_t(double v, int lane) { return readlane_d(v, lane); }
__device__ __forceinline__ float readlane_t(float v, int lane)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

// Panel factorization with the panel in registers (m <= 128 rows), wave 0:
// lane l holds rows l and l + 64 of the m x kb panel P.  Rows are not moved
// while the panel is factored: each register row carries its current row
// position pos (LAPACK's row order after the swaps so far), a swap exchanges
// two positions, and the rows are stored at their final positions at the
// end, so L and U come out in LAPACK's order.  The pivot is the first
// maximum |value| in row-position order (idamax), broadcast with v_readlane.
// b (nullable, LDS): the right-hand side, swapped and eliminated along.
// Returns the 1-based column of the first zero pivot in this panel, or 0.
template <typename T, typename TB>
__device__ __forceinline__ int panel_factor_regs(T *P, int ldP, int m, int kb, int *perm, TB *b)
{
    const int lane = threadIdx.x & 63;
    constexpr int kNone = 1 << 30;
    T v[2][kNb];
    TB bv[2];   // right-hand side entries of the register rows (they travel with the rows)
    int pos[2];
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        const int r = lane + 64 * q;
        pos[q] = r < m ? r : kNone;
#pragma unroll
        for (int c = 0; c < kNb; ++c) v[q][c] = (r < m && c < kb) ? P[(int64_t)r * ldP + c] : (T)0;
        bv[q] = (b && r < m) ? b[r] : (TB)0;
    }
    int info = 0;
    // fully unrolled over the panel's columns (jj a constant in every copy), so
    // the register rows are never indexed at run time; everything the column
    // steps touch is in registers, so they need no memory ordering
#pragma unroll
    for (int jj = 0; jj < kNb; ++jj) {
        if (jj < kb) {
            double best = -1.0;
            int bi = kNone;
#pragma unroll
            for (int q = 0; q < 2; ++q) {
                const double a = fabs((double)v[q][jj]);
                if (pos[q] >= jj && pos[q] < m && (a > best || (a == best && pos[q] < bi))) {
                    best = a;
                    bi = pos[q];
                }
            }
            const int pl = wave_argmax(best, bi);
            // the register row holding position pl: (lane lp, slot qp)
            const uint64_t own = __ballot(pos[0] == pl || pos[1] == pl);
            const int lp = __builtin_ctzll(own);
            const int qp = __builtin_amdgcn_readlane(pos[1] == pl ? 1 : 0, lp);
            const T pv = readlane_t(qp ? v[1][jj] : v[0][jj], lp);
            if (pv == (T)0) {
                if (info == 0) info = jj + 1;
                if (lane == 0) perm[jj] = jj;
            } else {
                if (lane == 0) perm[jj] = pl;
                if (pl != jj) {
#pragma unroll
                    for (int q = 0; q < 2; ++q)
                        pos[q] = pos[q] == pl ? jj : (pos[q] == jj ? pl : pos[q]);
                }
                // pivot row broadcast (qp is uniform: a scalar branch)
                T prow[kNb];
                TB bj;
                if (qp) {
#pragma unroll
                    for (int c = 0; c < kNb; ++c) prow[c] = c > jj ? readlane_t(v[1][c], lp) : (T)0;
                    bj = readlane_t(bv[1], lp);
                } else {
#pragma unroll
                    for (int c = 0; c < kNb; ++c) prow[c] = c > jj ? readlane_t(v[0][c], lp) : (T)0;
                    bj = readlane_t(bv[0], lp);
                }
                const T inv = (T)1 / pv;
                // rows below the pivot take l = v / pivot, the others l = 0 (an
                // exact no-op update), so the update is branch-free
#pragma unroll
                for (int q = 0; q < 2; ++q) {
                    const bool below = pos[q] > jj && pos[q] < m;
                    const T l = below ? v[q][jj] * inv : (T)0;
                    v[q][jj] = below ? l : v[q][jj];
#pragma unroll
                    for (int c = 0; c < kNb; ++c)
                        if (c > jj) v[q][c] = fma(-l, prow[c], v[q][c]);
                    bv[q] = below ? bv[q] - (TB)l * bj : bv[q];
                }
            }
        }
    }
#pragma unroll
    for (int q = 0; q < 2; ++q)
        if (pos[q] < m) {
#pragma unroll
            for (int c = 0; c < kNb; ++c)
                if (c < kb) P[(int64_t)pos[q] * ldP + c] = v[q][c];
            if (b) b[pos[q]] = bv[q];
        }
    wave_sync();
    return info;
}

typedef double fv_d4 __attribute__((ext_vector_type(4)));

// A22 -= L21 U12 (rows / columns k0+kb .. n, inner dimension kb <= 16), FP64
// on MFMA: each wavefront owns 16 x 16 tiles of A22 and accumulates them with
// v_mfma_f64_16x16x4f64 (A operand lane l: row l % 16, k = l / 16; B operand:
// column l % 16, k = l / 16; result element v of lane l: row 4 v + l / 16,
// column l % 16).  Out-of-range rows / columns / k enter as zeros.
__device__ __forceinline__ void trailing_update(double *a, int lda, int n, int k0, int kb)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    const int r0 = k0 + kb, mm = n - r0;
    const int tiles = (mm + 15) >> 4;
    const int lr = lane & 15, lk = lane >> 4;
    for (int t = wave; t < tiles * tiles; t += nw) {
        const int ti = t / tiles, tj = t - ti * tiles;
        const int R0 = r0 + 16 * ti, col = r0 + 16 * tj + lr;
        fv_d4 acc;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int row = R0 + 4 * v + lk;
            acc[v] = (row < n && col < n) ? a[(int64_t)row * lda + col] : 0.0;
        }
        const int ra = R0 + lr;
#pragma unroll
        for (int ks = 0; ks < kNb; ks += 4) {
            const int k = ks + lk;
            const double av = (ra < n && k < kb) ? a[(int64_t)ra * lda + k0 + k] : 0.0;
            const double bv = (col < n && k < kb) ? a[(int64_t)(k0 + k) * lda + col] : 0.0;
            acc = __builtin_amdgcn_mfma_f64_16x16x4f64(-av, bv, acc, 0, 0, 0);
        }
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int row = R0 + 4 * v + lk;
            if (row < n && col < n) a[(int64_t)row * lda + col] = acc[v];
        }
    }
}

typedef float fv_f4x __attribute__((ext_vector_type(4)));

// FP32 factors (mixed precision) on v_mfma_f32_16x16x4f32: operands as the
// FP64 form; result element v of lane l: row 4 (l / 16) + v, column l % 16.
__device__ __forceinline__ void trailing_update(float *a, int lda, int n, int k0, int kb)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    const int r0 = k0 + kb, mm = n - r0;
    const int tiles = (mm + 15) >> 4;
    const int lr = lane & 15, lk = lane >> 4;
    for (int t = wave; t < tiles * tiles; t += nw) {
        const int ti = t / tiles, tj = t - ti * tiles;
        const int R0 = r0 + 16 * ti, col = r0 + 16 * tj + lr;
        fv_f4x acc;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int row = R0 + 4 * lk + v;
            acc[v] = (row < n && col < n) ? a[(int64_t)row * lda + col] : 0.0f;
        }
        const int ra = R0 + lr;
#pragma unroll
        for (int ks = 0; ks < kNb; ks += 4) {
            const int k = ks + lk;
            const float av = (ra < n && k < kb) ? a[(int64_t)ra * lda + k0 + k] : 0.0f;
            const float bv = (col < n && k < kb) ? a[(int64_t)(k0 + k) * lda + col] : 0.0f;
            acc = __builtin_amdgcn_mfma_f32_16x16x4f32(-av, bv, acc, 0, 0, 0);
        }
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int row = R0 + 4 * lk + v;
            if (row < n && col < n) a[(int64_t)row * lda + col] = acc[v];
        }
    }
}

// Panel factorization staged through LDS (any m), wave 0: the m x kb panel P
// with row stride ldP, swapping the panel's rows and eliminating b as it goes.
template <typename T, typename TB>
__device__ __forceinline__ int panel_factor_lds(T *P, int ldP, int m, int kb, int *perm, TB *b)
{
    const int lane = threadIdx.x & 63;
    int info = 0;
    for (int jj = 0; jj < kb; ++jj) {
        double best = -1.0;
        int bi = m;
        for (int r = jj + lane; r < m; r += 64) {
            const double v = fabs((double)P[(int64_t)r * ldP + jj]);
            if (v > best) {
                best = v;
                bi = r;
            }
        }
        const int piv = wave_argmax(best, bi);
        if (P[(int64_t)piv * ldP + jj] == (T)0) {
            if (info == 0) info = jj + 1;
            if (lane == 0) perm[jj] = jj;
            wave_sync();
            continue;
        }
        if (lane == 0) perm[jj] = piv;
        if (piv != jj) {
            for (int c = lane; c < kb; c += 64) {
                const T t = P[(int64_t)jj * ldP + c];
                P[(int64_t)jj * ldP + c] = P[(int64_t)piv * ldP + c];
                P[(int64_t)piv * ldP + c] = t;
            }
            if (b && lane == 0) {
                const TB t = b[jj];
                b[jj] = b[piv];
                b[piv] = t;
            }
        }
        wave_sync();
        const T inv = (T)1 / P[(int64_t)jj * ldP + jj];
        const TB bj = b ? b[jj] : (TB)0;
        // pivot row in registers; each lane loads its row slice before
        // storing, so no LDS store waits on the next load
        T prow[kNb];
#pragma unroll
        for (int c = 0; c < kNb; ++c)
            prow[c] = (c > jj && c < kb) ? P[(int64_t)jj * ldP + c] : (T)0;
        for (int r = jj + 1 + lane; r < m; r += 64) {
            T *row = P + (int64_t)r * ldP;
            T v[kNb];
#pragma unroll
            for (int c = 0; c < kNb; ++c) v[c] = (c > jj && c < kb) ? row[c] : (T)0;
            const T l = row[jj] * inv;
            row[jj] = l;
#pragma unroll
            for (int c = 0; c < kNb; ++c)
                if (c > jj && c < kb) row[c] = v[c] - l * prow[c];
            if (b) b[r] -= (TB)l * bj;
        }
        wave_sync();
    }
    return info;
}

// a: n x n, row stride lda (LDS or global); pan: LDS panel buffer (n x kLdp)
// or null when a is in LDS; perm (LDS, n): perm[j] = row swapped with j;
// b (LDS, n, nullable): right-hand side, eliminated along.  Returns 0 or the
// 1-based index of the first zero pivot.  With REGS, panels of at most 128
// rows are factored in registers (panel_factor_regs: about 100 VGPRs, for
// kernels with that budget), taller ones through LDS.
template <typename T, typename TB, bool REGS = true>
__device__ __forceinline__ int lu_blocked(T *a, int lda, int n, int *perm, TB *b, T *pan)
{
    __shared__ int info_s;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    if (threadIdx.x == 0) info_s = 0;
    __syncthreads();
    for (int k0 = 0; k0 < n; k0 += kNb) {
        const int kb = (n - k0 < kNb) ? n - k0 : kNb;
        const int m = n - k0;
        T *P = a + (int64_t)k0 * lda + k0;
        const bool regs = REGS && m <= 128;
        if (!regs && pan) {
            for (int e = threadIdx.x; e < m * kb; e += blockDim.x) {
                const int r = e / kb, c = e - r * kb;
                pan[r * kLdp + c] = P[(int64_t)r * lda + c];
            }
            __syncthreads();
        }
        if (wave == 0) {
            TB *bk = b ? b + k0 : (TB *)nullptr;
            const int inf = regs ? panel_factor_regs<T, TB>(P, lda, m, kb, perm + k0, bk)
                            : pan ? panel_factor_lds<T, TB>(pan, kLdp, m, kb, perm + k0, bk)
                                  : panel_factor_lds<T, TB>(P, lda, m, kb, perm + k0, bk);
            if (lane == 0) {
                for (int jj = 0; jj < kb; ++jj) perm[k0 + jj] += k0;
                if (inf && info_s == 0) info_s = k0 + inf;
            }
        }
        __syncthreads();
        if (!regs && pan) {
            for (int e = threadIdx.x; e < m * kb; e += blockDim.x) {
                const int r = e / kb, c = e - r * kb;
                P[(int64_t)r * lda + c] = pan[r * kLdp + c];
            }
        }
        // the panel's row swaps on every column outside it, in pivot order
        for (int c = threadIdx.x; c < n - kb; c += blockDim.x) {
            const int cc = c < k0 ? c : c + kb;
            for (int jj = 0; jj < kb; ++jj) {
                const int p = perm[k0 + jj];
                if (p != k0 + jj) {
                    T *x = a + (int64_t)(k0 + jj) * lda + cc, *y = a + (int64_t)p * lda + cc;
                    const T t = *x;
                    *x = *y;
                    *y = t;
                }
            }
        }
        __syncthreads();
        // U12 = L11^{-1} A12: forward substitution, one column per lane in registers
        for (int c = k0 + kb + threadIdx.x; c < n; c += blockDim.x) {
            T col[kNb];
#pragma unroll
            for (int i = 0; i < kNb; ++i) col[i] = i < kb ? a[(int64_t)(k0 + i) * lda + c] : (T)0;
#pragma unroll
            for (int i = 1; i < kNb; ++i) {
                if (i < kb) {
                    T s = col[i];
#pragma unroll
                    for (int q = 0; q < i; ++q) s -= a[(int64_t)(k0 + i) * lda + k0 + q] * col[q];
                    col[i] = s;
                }
            }
#pragma unroll
            for (int i = 1; i < kNb; ++i)
                if (i < kb) a[(int64_t)(k0 + i) * lda + c] = col[i];
        }
        __syncthreads();
        // A22 -= L21 U12
        if (n - k0 - kb > 0) trailing_update(a, lda, n, k0, kb);
        __syncthreads();
    }
    return info_s;
}

// ------------------------------------------------ register-resident whole LU
// LU of an n x n matrix, n <= 128, by a workgroup of NW wavefronts (4, 8 or
// 16) with the whole matrix in registers: wavefront w holds columns w + NW mj
// (mj = 0 .. 128 / NW - 1), lane l rows l and l + 64.  Each column thus lives
// in one wavefront (its
// pivot search is a wavefront reduction, no barrier) and each row in lane
// l % 64 of every wavefront (the pivot row's entries reach each wavefront's
// lanes with v_readlane).  As in the register panels above, rows are not
// moved: every row carries its LAPACK row position (the same copy in every
// wavefront), a pivot exchanges two positions, and the rows are stored at
// their final positions at the end, so L and U come out in LAPACK's row
// order with LAPACK's pivot choices (first maximum in position order).
// One workgroup barrier per column: during step k the owner of column k+1
// updates that column first, finds its pivot and publishes the pivot row's
// lane / slot / position and every row's multiplier (double-buffered in LDS)
// while the other wavefronts finish the update of step k.
constexpr int kLuRegMax = 128;

template <typename T>
struct LuRegShared {
    T l[2][2][64];   // multipliers per buffer, register slot, lane
    int rec[2][4];   // per buffer: pivot lane, slot, position before the swap
    int info;        // 1-based column of the first exactly zero pivot, or 0
};

// max(a, b) of doubles without the NaN canonicalisation fmax() adds (the
// operands here are magnitudes, the sentinel -1 and zeros, never NaN).
__device__ __forceinline__ double dmax_raw(double a, double b)
{
    double r;
    asm("v_max_f64 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
    return r;
}

// DPP move of a double whose lanes without a source (or outside ROWMASK)
// read 0, the identity of a maximum over values >= 0.
template <int CTRL, int ROWMASK>
__device__ __forceinline__ double dpp_d0(double v)
{
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, ROWMASK, 0xf, true);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, ROWMASK, 0xf, true);
    return __hiloint2double(hi, lo);
}

// Largest value over the wavefront when at least one lane holds a value
// >= 0 (lanes without a candidate pass -1).
__device__ __forceinline__ double wave_max_nonneg(double m)
{
    m = dmax_raw(m, dpp_d0<0x111, 0xf>(m));
    m = dmax_raw(m, dpp_d0<0x112, 0xf>(m));
    m = dmax_raw(m, dpp_d0<0x114, 0xf>(m));
    m = dmax_raw(m, dpp_d0<0x118, 0xf>(m));
    m = dmax_raw(m, dpp_d0<0x142, 0xa>(m));
    m = dmax_raw(m, dpp_d0<0x143, 0xc>(m));
    return readlane_d(m, 63);
}

__device__ __forceinline__ int wave_min_i(int c)
{
    const int big = 0x7fffffff;
    c = min(c, dpp_i<0x111, 0xf>(big, c));
    c = min(c, dpp_i<0x112, 0xf>(big, c));
    c = min(c, dpp_i<0x114, 0xf>(big, c));
    c = min(c, dpp_i<0x118, 0xf>(big, c));
    c = min(c, dpp_i<0x142, 0xa>(big, c));
    c = min(c, dpp_i<0x143, 0xc>(big, c));
    return __builtin_amdgcn_readlane(c, 63);
}

// 1 / p for the multipliers l = v (1 / p), as LAPACK's dgetf2 scales the
// column by the reciprocal of the pivot: the hardware estimate refined by
// Newton steps (FP64: one step, measured max relative error 2.2e-15; FP32:
// one step) instead of the IEEE division sequence on the pivot chain.
__device__ __forceinline__ double lureg_recip(double p) { return rcp_refined(p); }
__device__ __forceinline__ float lureg_recip(float p)
{
    const float y = __builtin_amdgcn_rcpf(p);
    return fmaf(y, fmaf(-p, y, 1.0f), y);
}

// Pivot of column jj (register column MJ of the calling wavefront, its
// owner) among the rows at positions >= jj: the largest |value|, ties to
// the smallest position (idamax).  One maximum reduction; the lanes holding
// that maximum are found by a ballot, and only an exact tie between lanes
// takes a second (position) reduction.  Stores the multipliers into the
// column and publishes the step record into buffer buf.
template <int MJ, int CW, typename T>
__device__ __forceinline__ void lureg_pivot(T (&v)[2][CW], const int (&pos)[2], int jj, int n,
                                            LuRegShared<T> *sh, int buf)
{
    const int lane = threadIdx.x & 63;
    // the lane's candidate: the larger |value| of its two rows, ties to the
    // smaller position (branch-free; non-candidates enter as -1)
    const double b0 = (pos[0] >= jj && pos[0] < n) ? fabs((double)v[0][MJ]) : -1.0;
    const double b1 = (pos[1] >= jj && pos[1] < n) ? fabs((double)v[1][MJ]) : -1.0;
    const bool t1 = b1 > b0 || (b1 == b0 && pos[1] < pos[0]);
    const double best = t1 ? b1 : b0;
    const int bq = t1 ? 1 : 0;
    const int bp = t1 ? pos[1] : pos[0];
    const double M = wave_max_nonneg(best);
    const uint64_t hit = __ballot(best == M);
    int lp;
    if (__popcll(hit) == 1) {
        lp = __builtin_ctzll(hit);
    } else {
        const int pmin = wave_min_i(best == M ? bp : 0x7fffffff);
        lp = __builtin_ctzll(__ballot(best == M && bp == pmin));
    }
    const int qp = __builtin_amdgcn_readlane(bq, lp);
    const int pl = __builtin_amdgcn_readlane(bp, lp);
    const T pv = readlane_t(qp ? v[1][MJ] : v[0][MJ], lp);
    const bool zero = pv == (T)0;
    const T inv = zero ? (T)0 : lureg_recip(pv);
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        // rows below the pivot after its exchange take l = v / pivot, the
        // others l = 0 (an exact no-op update)
        const bool below = !zero && pos[q] >= jj && pos[q] < n && pos[q] != pl;
        const T l = below ? v[q][MJ] * inv : (T)0;
        v[q][MJ] = below ? l : v[q][MJ];
        sh->l[buf][q][lane] = l;
    }
    if (lane == 0) {
        sh->rec[buf][0] = lp;
        sh->rec[buf][1] = qp;
        sh->rec[buf][2] = pl;
        if (zero && sh->info == 0) sh->info = jj + 1;
    }
}

// v[q][mj] -= l[q] * (pivot row entry of column wv + NW mj), QP the pivot
// row's register slot: for every mj > MJ0 (those columns lie beyond step k
// of block MJ0; padding columns >= n take the update too, harmlessly, so the
// loop has no branches) and, when first, for mj == MJ0.
template <int MJ0, int QP, int CW, typename T>
__device__ __forceinline__ void lureg_update(T (&v)[2][CW], const T (&l)[2], int lp, bool first)
{
    if (first) {
        const T u = readlane_t(v[QP][MJ0], lp);
#pragma unroll
        for (int q = 0; q < 2; ++q) v[q][MJ0] = fma(-l[q], u, v[q][MJ0]);
    }
#pragma unroll
    for (int mj = MJ0 + 1; mj < CW; ++mj) {
        const T u = readlane_t(v[QP][mj], lp);
#pragma unroll
        for (int q = 0; q < 2; ++q) v[q][mj] = fma(-l[q], u, v[q][mj]);
    }
}

// The previous step's multipliers and pivot row, kept by every wavefront:
// the owner of the look-ahead column applies the rest of that step's update
// one step late (after publishing its pivot), so the barrier does not wait
// for it.
template <typename T>
struct LuRegPrev {
    T l[2];
    int lp, qp;
};

// Step k of the elimination (k / NW == MJ0) by every wavefront (wv uniform).
template <int MJ0, int NW, typename T>
__device__ __forceinline__ void lureg_step(T (&v)[2][kLuRegMax / NW], int (&pos)[2], int k,
                                           int n, int wv, LuRegShared<T> *sh, LuRegPrev<T> &prev)
{
    constexpr int CW = kLuRegMax / NW;
    const int lane = threadIdx.x & 63;
    const int buf = k & 1;
    const int lp = __builtin_amdgcn_readfirstlane(sh->rec[buf][0]);
    const int qp = __builtin_amdgcn_readfirstlane(sh->rec[buf][1]);
    const int pl = __builtin_amdgcn_readfirstlane(sh->rec[buf][2]);
    const T l[2] = {sh->l[buf][0][lane], sh->l[buf][1][lane]};
#pragma unroll
    for (int q = 0; q < 2; ++q) pos[q] = pos[q] == pl ? k : (pos[q] == k ? pl : pos[q]);
    const int kl = k & (NW - 1);
    // the owner of column k (the previous step's look-ahead) catches up on
    // the previous step's update of its later columns
    if (k > 0 && kl == wv) {
        if (prev.qp)
            lureg_update<MJ0, 1>(v, prev.l, prev.lp, false);
        else
            lureg_update<MJ0, 0>(v, prev.l, prev.lp, false);
    }
    const int k1 = k + 1;
    if (k1 < n && (k1 & (NW - 1)) == wv) {
        // look-ahead: column k+1 first, then its pivot; the rest of this
        // step's update follows in the next step.  This wavefront is the
        // step's critical path: it issues ahead of the others on its SIMD
        // (wave priority) until the pivot is published.
        __builtin_amdgcn_s_setprio(3);
        if (kl != NW - 1) {
            const T u = readlane_t(qp ? v[1][MJ0] : v[0][MJ0], lp);
#pragma unroll
            for (int q = 0; q < 2; ++q) v[q][MJ0] = fma(-l[q], u, v[q][MJ0]);
            lureg_pivot<MJ0>(v, pos, k1, n, sh, buf ^ 1);
        } else if (MJ0 < CW - 1) {
            constexpr int M1 = MJ0 < CW - 1 ? MJ0 + 1 : CW - 1;
            const T u = readlane_t(qp ? v[1][M1] : v[0][M1], lp);
#pragma unroll
            for (int q = 0; q < 2; ++q) v[q][M1] = fma(-l[q], u, v[q][M1]);
            lureg_pivot<M1>(v, pos, k1, n, sh, buf ^ 1);
        }
        __builtin_amdgcn_s_setprio(0);
    } else if (qp) {
        lureg_update<MJ0, 1>(v, l, lp, wv > kl);
    } else {
        lureg_update<MJ0, 0>(v, l, lp, wv > kl);
    }
    prev.l[0] = l[0];
    prev.l[1] = l[1];
    prev.lp = lp;
    prev.qp = qp;
    __syncthreads();
}

// Steps NW MJ0 .. min(n, NW MJ0 + NW) - 1, then the next block: one loop per
// block, so every register column index is a constant.
template <int MJ0, int NW, typename T>
__device__ __forceinline__ void lureg_block(T (&v)[2][kLuRegMax / NW], int (&pos)[2], int n,
                                            int wv, LuRegShared<T> *sh, LuRegPrev<T> &prev)
{
    const int kend = n < NW * MJ0 + NW ? n : NW * MJ0 + NW;
    for (int k = NW * MJ0; k < kend; ++k) lureg_step<MJ0, NW>(v, pos, k, n, wv, sh, prev);
    if constexpr (MJ0 < kLuRegMax / NW - 1) {
        if (n > NW * MJ0 + NW) lureg_block<MJ0 + 1, NW>(v, pos, n, wv, sh, prev);
    }
}

// Factors the n x n matrix a (row stride lda, n <= 128; LDS or global) by a
// workgroup of 64 NW lanes, in place: on return a holds L and U in
// LAPACK's row order and order[i] the original row of factor row i
// ((P v)[i] = v[order[i]]).  Returns the 1-based column of the first exactly
// zero pivot (the factorization continues past it, as LAPACK's), or 0.
template <int NW, typename T>
__device__ __forceinline__ int lu_regs(T *a, int lda, int n, int *order, LuRegShared<T> *sh)
{
    static_assert(NW == 4 || NW == 8 || NW == 16, "4, 8 or 16 wavefronts");
    constexpr int CW = kLuRegMax / NW;
    const int lane = threadIdx.x & 63;
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    constexpr int kNone = 1 << 30;
    T v[2][CW];
    int pos[2];
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        const int r = lane + 64 * q;
        pos[q] = r < n ? r : kNone;
#pragma unroll
        for (int mj = 0; mj < CW; ++mj) {
            const int j = wv + NW * mj;
            v[q][mj] = (r < n && j < n) ? a[r * lda + j] : (T)0;
        }
    }
    if (threadIdx.x == 0) sh->info = 0;
    if (wv == 0) lureg_pivot<0>(v, pos, 0, n, sh, 0);
    __syncthreads();
    LuRegPrev<T> prev{{(T)0, (T)0}, 0, 0};
    lureg_block<0, NW>(v, pos, n, wv, sh, prev);
    // every row to its final position (a was read before the first barrier)
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        if (pos[q] < n) {
#pragma unroll
            for (int mj = 0; mj < CW; ++mj) {
                const int j = wv + NW * mj;
                if (j < n) a[pos[q] * lda + j] = v[q][mj];
            }
            if (wv == 0) order[pos[q]] = lane + 64 * q;
        }
    }
    const int inf = sh->info;
    __syncthreads();
    return inf;
}

// Register form of the substitutions (n <= 256): lane l holds x[l + 64 q].
// Blocks of 16 columns: the block's own rows are solved sequentially from
// the 16 x 16 diagonal block held in registers (x broadcast with
// v_readlane), then every row outside the block is updated at once with
// the block's 16 finished values (independent LDS loads, no chain).
template <typename TB>
__device__ __forceinline__ TB slot_get(const TB *xr, int q)
{
    return q == 0 ? xr[0] : q == 1 ? xr[1] : q == 2 ? xr[2] : xr[3];
}

template <typename TB>
__device__ __forceinline__ void slot_set(TB *xr, int q, TB v)
{
#pragma unroll
    for (int k = 0; k < 4; ++k)
        if (k == q) xr[k] = v;
}

// x = U^{-1} x in place by wave 0 (U upper triangular, 1 / U_kk in rdiag).
// x lives in registers for n <= 256 (blocked, above), else in LDS
// (column-oriented, one broadcast per column).
template <typename T, typename TB>
__device__ __forceinline__ void wave_back_subst(const T *a, int lda, int n, TB *x,
                                                const TB *rdiag)
{
    if ((threadIdx.x >> 6) != 0) return;
    const int lane = threadIdx.x & 63;
    if (n <= 256) {
        TB xr[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) xr[q] = (lane + 64 * q < n) ? x[lane + 64 * q] : (TB)0;
        for (int kb0 = ((n - 1) / kNb) * kNb; kb0 >= 0; kb0 -= kNb) {
            const int kbl = (n - kb0 < kNb) ? n - kb0 : kNb;
            const int q0 = kb0 >> 6, l0 = kb0 & 63, ib = lane - l0;
            const bool inblk = ib >= 0 && ib < kbl;
            T u[kNb];
#pragma unroll
            for (int j = 0; j < kNb; ++j)
                u[j] = (inblk && j > ib && j < kbl) ? a[(int64_t)(kb0 + ib) * lda + kb0 + j] : (T)0;
            const TB rd = inblk ? rdiag[kb0 + ib] : (TB)0;
            TB xb = slot_get(xr, q0);
#pragma unroll
            for (int j = kNb - 1; j >= 0; --j) {
                if (j < kbl) {
                    const TB xj = readlane_t(xb, l0 + j) * readlane_t(rd, l0 + j);
                    xb = ib == j ? xj : fma(-(TB)u[j], xj, xb);   // u[j] = 0 off the block's upper part
                }
            }
            slot_set(xr, q0, xb);
            TB xs[kNb];
#pragma unroll
            for (int j = 0; j < kNb; ++j) xs[j] = j < kbl ? readlane_t(xb, l0 + j) : (TB)0;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int i = lane + 64 * q;
                if (i < kb0) {
                    TB acc = xr[q];
#pragma unroll
                    for (int j = 0; j < kNb; ++j)
                        if (j < kbl) acc -= (TB)a[(int64_t)i * lda + kb0 + j] * xs[j];
                    xr[q] = acc;
                }
            }
        }
#pragma unroll
        for (int q = 0; q < 4; ++q)
            if (lane + 64 * q < n) x[lane + 64 * q] = xr[q];
        wave_sync();
        return;
    }
    for (int k = n - 1; k >= 0; --k) {
        const TB xk = x[k] * rdiag[k];
        for (int i = lane; i < k; i += 64) x[i] -= (TB)a[(int64_t)i * lda + k] * xk;
        wave_sync();
        if (lane == 0) x[k] = xk;
        wave_sync();
    }
}

// y = L^{-1} x in place (unit lower triangle) by wave 0, blocked in registers
// for n <= 256 as above (REGS; about 80 VGPRs), else column-oriented in LDS;
// the row permutation is applied by the caller.
template <typename T, typename TB, bool REGS = true>
__device__ __forceinline__ void wave_fwd_subst(const T *a, int lda, int n, TB *x)
{
    if ((threadIdx.x >> 6) != 0) return;
    const int lane = threadIdx.x & 63;
    if (REGS && n <= 256) {
        TB xr[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) xr[q] = (lane + 64 * q < n) ? x[lane + 64 * q] : (TB)0;
        for (int kb0 = 0; kb0 < n; kb0 += kNb) {
            const int kbl = (n - kb0 < kNb) ? n - kb0 : kNb;
            const int q0 = kb0 >> 6, l0 = kb0 & 63, ib = lane - l0;
            const bool inblk = ib >= 0 && ib < kbl;
            T lrow[kNb];
#pragma unroll
            for (int j = 0; j < kNb; ++j)
                lrow[j] = (inblk && j < ib) ? a[(int64_t)(kb0 + ib) * lda + kb0 + j] : (T)0;
            TB xb = slot_get(xr, q0);
#pragma unroll
            for (int j = 0; j < kNb; ++j) {
                if (j < kbl) xb = fma(-(TB)lrow[j], readlane_t(xb, l0 + j), xb);   // lrow[j] = 0 off L
            }
            slot_set(xr, q0, xb);
            TB xs[kNb];
#pragma unroll
            for (int j = 0; j < kNb; ++j) xs[j] = j < kbl ? readlane_t(xb, l0 + j) : (TB)0;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int i = lane + 64 * q;
                if (i >= kb0 + kbl && i < n) {
                    TB acc = xr[q];
#pragma unroll
                    for (int j = 0; j < kNb; ++j)
                        if (j < kbl) acc -= (TB)a[(int64_t)i * lda + kb0 + j] * xs[j];
                    xr[q] = acc;
                }
            }
        }
#pragma unroll
        for (int q = 0; q < 4; ++q)
            if (lane + 64 * q < n) x[lane + 64 * q] = xr[q];
        wave_sync();
        return;
    }
    for (int k = 0; k < n - 1; ++k) {
        const TB xk = x[k];
        for (int i = k + 1 + lane; i < n; i += 64) x[i] -= (TB)a[(int64_t)i * lda + k] * xk;
        wave_sync();
    }
}

// Substitutions for n <= 128 by wave 0, x[i] in lane i % 64, slot i / 64 of
// registers: column-oriented, so the dependency chain per column is one
// broadcast (v_readlane of the finished entry) and one FMA; the column's
// coefficients are loaded from LDS in groups of 8 columns ahead of the chain
// and masked off it (rows the column must not change take 0, an exact no-op
// apart from the sign of a zero).  Loads of padding rows (>= n) read row
// n - 1; their results are never stored.
//   forward:  y = L^{-1} x (unit lower triangle)
//   backward: x = U^{-1} y with y_k scaled by rdiag[k] = 1 / U_kk; the
//             update of row i by column k takes (U_ik rdiag[k]) y_k, and
//             x_i = y_i rdiag[i] at the end.
template <typename T, typename TB>
__device__ __forceinline__ void wave_fwd_subst128(const T *a, int lda, int n, TB *x)
{
    if ((threadIdx.x >> 6) != 0) return;
    const int lane = threadIdx.x & 63;
    const T *c0 = a + (lane < n ? lane : n - 1) * lda;
    const T *c1 = a + (lane + 64 < n ? lane + 64 : n - 1) * lda;
    TB x0 = lane < n ? x[lane] : (TB)0;
    TB x1 = lane + 64 < n ? x[lane + 64] : (TB)0;
    const int last = n - 1;                      // columns 0 .. n-2 update
    const int kA = last < 64 ? last : 64;        // columns whose entry sits in slot 0
    int k = 0;
    for (; k + 8 <= kA; k += 8) {
        TB l0[8], l1[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            l0[j] = lane > k + j ? (TB)c0[k + j] : (TB)0;
            l1[j] = (TB)c1[k + j];
        }
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const TB xk = readlane_t(x0, k + j);
            x0 = fma(-l0[j], xk, x0);
            x1 = fma(-l1[j], xk, x1);
        }
    }
    for (; k < kA; ++k) {
        const TB l0 = lane > k ? (TB)c0[k] : (TB)0, l1 = (TB)c1[k];
        const TB xk = readlane_t(x0, k);
        x0 = fma(-l0, xk, x0);
        x1 = fma(-l1, xk, x1);
    }
    for (; k + 8 <= last; k += 8) {
        TB l1[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) l1[j] = lane + 64 > k + j ? (TB)c1[k + j] : (TB)0;
#pragma unroll
        for (int j = 0; j < 8; ++j) x1 = fma(-l1[j], readlane_t(x1, k + j - 64), x1);
    }
    for (; k < last; ++k) {
        const TB l1 = lane + 64 > k ? (TB)c1[k] : (TB)0;
        x1 = fma(-l1, readlane_t(x1, k - 64), x1);
    }
    if (lane < n) x[lane] = x0;
    if (lane + 64 < n) x[lane + 64] = x1;
    wave_sync();
}

template <typename T, typename TB>
__device__ __forceinline__ void wave_back_subst128(const T *a, int lda, int n, TB *x,
                                                   const TB *rdiag)
{
    if ((threadIdx.x >> 6) != 0) return;
    const int lane = threadIdx.x & 63;
    const T *c0 = a + (lane < n ? lane : n - 1) * lda;
    const T *c1 = a + (lane + 64 < n ? lane + 64 : n - 1) * lda;
    TB x0 = lane < n ? x[lane] : (TB)0;
    TB x1 = lane + 64 < n ? x[lane + 64] : (TB)0;
    // columns n-1 .. 64 (entry in slot 1): rows above in both slots
    int k = n - 1;
    for (; k - 8 >= 63; k -= 8) {
        TB u0[8], u1[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const TB rd = rdiag[k - j];
            u0[j] = (TB)c0[k - j] * rd;
            u1[j] = lane + 64 < k - j ? (TB)c1[k - j] * rd : (TB)0;
        }
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const TB yk = readlane_t(x1, k - j - 64);
            x0 = fma(-u0[j], yk, x0);
            x1 = fma(-u1[j], yk, x1);
        }
    }
    for (; k >= 64; --k) {
        const TB rd = rdiag[k];
        const TB u0 = (TB)c0[k] * rd, u1 = lane + 64 < k ? (TB)c1[k] * rd : (TB)0;
        const TB yk = readlane_t(x1, k - 64);
        x0 = fma(-u0, yk, x0);
        x1 = fma(-u1, yk, x1);
    }
    // columns min(n-1, 63) .. 1 (entry in slot 0): rows above in slot 0
    for (; k - 8 >= 0; k -= 8) {
        TB u0[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) u0[j] = lane < k - j ? (TB)c0[k - j] * rdiag[k - j] : (TB)0;
#pragma unroll
        for (int j = 0; j < 8; ++j) x0 = fma(-u0[j], readlane_t(x0, k - j), x0);
    }
    for (; k >= 1; --k) {
        const TB u0 = lane < k ? (TB)c0[k] * rdiag[k] : (TB)0;
        x0 = fma(-u0, readlane_t(x0, k), x0);
    }
    if (lane < n) x[lane] = x0 * rdiag[lane];
    if (lane + 64 < n) x[lane + 64] = x1 * rdiag[lane + 64];
    wave_sync();
}

}  // namespace fv
