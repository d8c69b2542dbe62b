// fv_linalg.hip -- batched dense circulation solves.
//
// One workgroup per system (the reference's batched solver runs one thread
// per system, warp_fsi/batched_solver.py:20-63).  Blocked LU with partial
// pivoting (first maximum wins, as LAPACK's idamax), the matrix held in LDS
// when it fits and in global memory (L2-resident at these sizes) otherwise.
//   FP64:  factor + solve in FP64 (np.linalg.solve, standalone_uvlm.py:427).
//   Mixed: factor in FP32, then iterative refinement with FP64 residuals
//          r = b - A x against the FP64 matrix.
#include <cmath>

#include "fv_launch.h"

namespace fv {

constexpr int kLuBlock = 256;

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

// a: n x n, row stride lda (LDS or global); pan: LDS panel buffer (n x kLdp)
// or null when a is in LDS; perm (LDS, n): perm[j] = row swapped with j;
// b (LDS, n, nullable): right-hand side, eliminated along.  Returns 0 or the
// 1-based index of the first zero pivot.
template <typename T, typename TB>
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
        int ldP = lda;
        if (pan) {
            for (int e = threadIdx.x; e < m * kb; e += blockDim.x) {
                const int r = e / kb, c = e - r * kb;
                pan[r * kLdp + c] = P[(int64_t)r * lda + c];
            }
            P = pan;
            ldP = kLdp;
            __syncthreads();
        }
        if (wave == 0) {
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
                    if (lane == 0) {
                        if (info_s == 0) info_s = k0 + jj + 1;
                        perm[k0 + jj] = k0 + jj;
                    }
                    wave_sync();
                    continue;
                }
                if (lane == 0) perm[k0 + jj] = k0 + piv;
                if (piv != jj) {
                    for (int c = lane; c < kb; c += 64) {
                        const T t = P[(int64_t)jj * ldP + c];
                        P[(int64_t)jj * ldP + c] = P[(int64_t)piv * ldP + c];
                        P[(int64_t)piv * ldP + c] = t;
                    }
                    if (b && lane == 0) {
                        const TB t = b[k0 + jj];
                        b[k0 + jj] = b[k0 + piv];
                        b[k0 + piv] = t;
                    }
                }
                wave_sync();
                const T inv = (T)1 / P[(int64_t)jj * ldP + jj];
                const TB bj = b ? b[k0 + jj] : (TB)0;
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
                    if (b) b[k0 + r] -= (TB)l * bj;
                }
                wave_sync();
            }
        }
        __syncthreads();
        if (pan) {
            T *G = a + (int64_t)k0 * lda + k0;
            for (int e = threadIdx.x; e < m * kb; e += blockDim.x) {
                const int r = e / kb, c = e - r * kb;
                G[(int64_t)r * lda + c] = pan[r * kLdp + c];
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

// x = U^{-1} x in place by wave 0, column-oriented: x[k] is final once the
// columns right of k have been eliminated; it is broadcast with v_readlane
// and column k is eliminated from the rows above.  x lives in registers
// (lane l holds x[l + 64 q]) for n <= 256, else in LDS.
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
        for (int k = n - 1; k >= 0; --k) {
            const int ks = k >> 6, kl = k & 63;
            const TB own = ks == 0 ? xr[0] : ks == 1 ? xr[1] : ks == 2 ? xr[2] : xr[3];
            const TB xk = readlane_d((double)own, kl) * rdiag[k];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int i = lane + 64 * q;
                if (i < k) xr[q] -= (TB)a[(int64_t)i * lda + k] * xk;
                if (i == k) xr[q] = xk;
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

// y = L^{-1} x in place (unit lower triangle) by wave 0, column-oriented as
// above; the row permutation is applied by the caller.
template <typename T, typename TB>
__device__ __forceinline__ void wave_fwd_subst(const T *a, int lda, int n, TB *x)
{
    if ((threadIdx.x >> 6) != 0) return;
    const int lane = threadIdx.x & 63;
    if (n <= 256) {
        TB xr[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) xr[q] = (lane + 64 * q < n) ? x[lane + 64 * q] : (TB)0;
        for (int k = 0; k < n - 1; ++k) {
            const int ks = k >> 6, kl = k & 63;
            const TB own = ks == 0 ? xr[0] : ks == 1 ? xr[1] : ks == 2 ? xr[2] : xr[3];
            const TB xk = readlane_d((double)own, kl);
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int i = lane + 64 * q;
                if (i > k && i < n) xr[q] -= (TB)a[(int64_t)i * lda + k] * xk;
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

// LDS plan of one solve: the whole matrix (odd row stride) when it fits,
// else a staged panel; then the vectors.
struct LuLds {
    int in_lds;      // matrix in LDS
    int lda;         // row stride of the matrix
    size_t mat;      // bytes of the matrix or panel area
    size_t total;    // dynamic LDS bytes
};

__host__ __device__ inline size_t lu_align(size_t v) { return (v + 15) & ~size_t(15); }

template <typename T>
static LuLds lu_lds_plan(int64_t n, size_t vec_bytes, size_t limit)
{
    LuLds p;
    const int64_t ldo = n | 1;
    const size_t full = lu_align((size_t)n * ldo * sizeof(T));
    if (full + vec_bytes <= limit) {
        p.in_lds = 1;
        p.lda = (int)ldo;
        p.mat = full;
    } else {
        p.in_lds = 0;
        p.lda = (int)n;
        p.mat = lu_align((size_t)n * kLdp * sizeof(T));
    }
    p.total = p.mat + vec_bytes;
    return p;
}

// IN_LDS: the matrix is copied into LDS (row stride lda_lds) and factored
// there; else it is factored in global memory with the panels staged in LDS.
// (A compile-time choice, so LDS accesses compile to ds_* instructions.)
template <int IN_LDS>
__global__ __launch_bounds__(kLuBlock) void k_lu_solve_f64(int n, double *__restrict__ A,
                                                           double *__restrict__ B,
                                                           int32_t *__restrict__ info,
                                                           int lda_lds, unsigned mat_bytes)
{
    extern __shared__ __attribute__((aligned(16))) unsigned char lds_raw[];
    double *a = A + (int64_t)blockIdx.x * n * n;
    double *bg = B + (int64_t)blockIdx.x * n;
    double *area = reinterpret_cast<double *>(lds_raw);
    double *b = reinterpret_cast<double *>(lds_raw + mat_bytes);
    double *rdiag = b + n;
    int *perm = reinterpret_cast<int *>(rdiag + n);
    double *w = IN_LDS ? area : a;
    const int lda = IN_LDS ? lda_lds : n;
    if (IN_LDS) {
        for (int64_t e = threadIdx.x; e < (int64_t)n * n; e += kLuBlock) {
            const int r = (int)(e / n), c = (int)(e - (int64_t)r * n);
            w[(int64_t)r * lda + c] = a[e];
        }
    }
    for (int i = threadIdx.x; i < n; i += kLuBlock) b[i] = bg[i];
    __syncthreads();
    const int inf = lu_blocked<double, double>(w, lda, n, perm, b, IN_LDS ? nullptr : area);
    for (int i = threadIdx.x; i < n; i += kLuBlock) rdiag[i] = 1.0 / w[(int64_t)i * lda + i];
    __syncthreads();
    if (inf == 0) wave_back_subst<double, double>(w, lda, n, b, rdiag);
    __syncthreads();
    if (IN_LDS)
        for (int64_t e = threadIdx.x; e < (int64_t)n * n; e += kLuBlock) {
            const int r = (int)(e / n), c = (int)(e - (int64_t)r * n);
            a[e] = w[(int64_t)r * lda + c];
        }
    for (int i = threadIdx.x; i < n; i += kLuBlock) bg[i] = b[i];
    if (threadIdx.x == 0 && info) info[blockIdx.x] = inf;
}

// Mixed: FP32 factors (LDS or the global scratch f32), FP64 refinement with
// the vectors in LDS.
template <int IN_LDS>
__global__ __launch_bounds__(kLuBlock) void k_lu_solve_mixed(
    int n, const double *__restrict__ A, double *__restrict__ B, float *__restrict__ f32,
    int max_iter, double tol, int32_t *__restrict__ info, int32_t *__restrict__ iters,
    int lda_lds, unsigned mat_bytes)
{
    extern __shared__ __attribute__((aligned(16))) unsigned char lds_raw[];
    const double *a = A + (int64_t)blockIdx.x * n * n;
    double *bg = B + (int64_t)blockIdx.x * n;
    float *area = reinterpret_cast<float *>(lds_raw);
    double *x = reinterpret_cast<double *>(lds_raw + mat_bytes);
    double *r = x + n;
    double *d = r + n;
    double *rdiag = d + n;
    int *perm = reinterpret_cast<int *>(rdiag + n);
    int *order = perm + n;
    float *lu = IN_LDS ? area : f32 + (int64_t)blockIdx.x * n * n;
    const int lda = IN_LDS ? lda_lds : n;
    __shared__ double stat[2];
    for (int64_t e = threadIdx.x; e < (int64_t)n * n; e += kLuBlock) {
        const int rr = (int)(e / n), c = (int)(e - (int64_t)rr * n);
        lu[(int64_t)rr * lda + c] = (float)a[e];
    }
    __syncthreads();
    const int inf = lu_blocked<float, double>(lu, lda, n, perm, (double *)nullptr,
                                              IN_LDS ? nullptr : area);
    if (inf != 0) {
        if (threadIdx.x == 0) {
            if (info) info[blockIdx.x] = inf;
            if (iters) iters[blockIdx.x] = 0;
        }
        return;
    }
    for (int i = threadIdx.x; i < n; i += kLuBlock) {
        x[i] = 0.0;
        r[i] = bg[i];
        order[i] = i;
        rdiag[i] = 1.0 / (double)lu[(int64_t)i * lda + i];
    }
    __syncthreads();
    // the pivot sequence composed into one row order: (P v)[i] = v[order[i]]
    if (threadIdx.x == 0)
        for (int k = 0; k < n; ++k) {
            const int q = perm[k];
            if (q != k) {
                const int t = order[k];
                order[k] = order[q];
                order[q] = t;
            }
        }
    __syncthreads();
    int it = 0;
    for (; it <= max_iter; ++it) {
        // correction d = U32^{-1} L32^{-1} P r (carried in FP64 storage)
        for (int i = threadIdx.x; i < n; i += kLuBlock) d[i] = r[order[i]];
        __syncthreads();
        wave_fwd_subst<float, double>(lu, lda, n, d);
        wave_back_subst<float, double>(lu, lda, n, d, rdiag);
        __syncthreads();
        double dmax = 0.0, xmax = 0.0;
        if (threadIdx.x < 64) {
            for (int i = threadIdx.x; i < n; i += 64) {
                x[i] += d[i];
                dmax = fmax(dmax, fabs(d[i]));
                xmax = fmax(xmax, fabs(x[i]));
            }
            for (int off = 32; off > 0; off >>= 1) {
                dmax = fmax(dmax, __shfl_xor(dmax, off, 64));
                xmax = fmax(xmax, __shfl_xor(xmax, off, 64));
            }
            if (threadIdx.x == 0) {
                stat[0] = dmax;
                stat[1] = xmax;
            }
        }
        __syncthreads();
        const bool done = stat[0] <= tol * stat[1];
        if (done || it == max_iter) break;
        // FP64 residual r = b - A x: one wavefront per row, coalesced row reads,
        // DPP sum (row_shr 1/2/4/8, row_bcast 15/31; lane 63 holds the sum)
        for (int i = threadIdx.x >> 6; i < n; i += kLuBlock >> 6) {
            const int lane = threadIdx.x & 63;
            const double *row = a + (int64_t)i * n;
            double s = 0.0;
            for (int c = lane; c < n; c += 64) s = fma(row[c], x[c], s);
            s += dpp_d<0x111, 0xf>(0.0, s);
            s += dpp_d<0x112, 0xf>(0.0, s);
            s += dpp_d<0x114, 0xf>(0.0, s);
            s += dpp_d<0x118, 0xf>(0.0, s);
            s += dpp_d<0x142, 0xa>(0.0, s);
            s += dpp_d<0x143, 0xc>(0.0, s);
            const double tot = readlane_d(s, 63);
            if (lane == 0) r[i] = bg[i] - tot;
        }
        __syncthreads();
    }
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += kLuBlock) bg[i] = x[i];
    if (threadIdx.x == 0) {
        if (info) info[blockIdx.x] = 0;
        if (iters) iters[blockIdx.x] = it;
    }
}

static size_t max_lds(fv_ctx *ctx)
{
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, ctx->device) !=
            hipSuccess ||
        v <= 0)
        v = 65536;
    return (size_t)v;
}

int dense_solve(fv_ctx *ctx, int64_t batch, int64_t n, double *A, double *rhs, int32_t *info)
{
    if (batch <= 0 || n <= 0) return FV_OK;
    const size_t vec = 2 * lu_align((size_t)n * sizeof(double)) + lu_align((size_t)n * sizeof(int));
    const LuLds p = lu_lds_plan<double>(n, vec, max_lds(ctx));
    FV_REQUIRE(ctx, p.total <= max_lds(ctx), "dense solve: n = %lld exceeds the LDS panel limit",
               (long long)n);
    TimedLaunch tl(ctx, "dense_solve");
    if (p.in_lds)
        hipLaunchKernelGGL(k_lu_solve_f64<1>, dim3((unsigned)batch), dim3(kLuBlock),
                           (unsigned)p.total, ctx->stream, (int)n, A, rhs, info, p.lda,
                           (unsigned)p.mat);
    else
        hipLaunchKernelGGL(k_lu_solve_f64<0>, dim3((unsigned)batch), dim3(kLuBlock),
                           (unsigned)p.total, ctx->stream, (int)n, A, rhs, info, p.lda,
                           (unsigned)p.mat);
    FV_LAUNCH_CHECK(ctx);
    tl.done();
    return FV_OK;
}

int dense_solve_mixed(fv_ctx *ctx, int64_t batch, int64_t n, const double *A, double *rhs,
                      int max_iter, double tol, int32_t *info, int32_t *iters)
{
    if (batch <= 0 || n <= 0) return FV_OK;
    const size_t vec = 4 * lu_align((size_t)n * sizeof(double)) +
                       2 * lu_align((size_t)n * sizeof(int));
    const LuLds p = lu_lds_plan<float>(n, vec, max_lds(ctx));
    FV_REQUIRE(ctx, p.total <= max_lds(ctx), "dense solve: n = %lld exceeds the LDS panel limit",
               (long long)n);
    float *f32 = nullptr;
    if (!p.in_lds) {
        f32 = static_cast<float *>(
            workspace(ctx, S_MAT32, sizeof(float) * (size_t)batch * n * n));
        if (!f32) return FV_E_HIP;
    }
    TimedLaunch tl(ctx, "dense_solve");
    if (p.in_lds)
        hipLaunchKernelGGL(k_lu_solve_mixed<1>, dim3((unsigned)batch), dim3(kLuBlock),
                           (unsigned)p.total, ctx->stream, (int)n, A, rhs, f32, max_iter, tol,
                           info, iters, p.lda, (unsigned)p.mat);
    else
        hipLaunchKernelGGL(k_lu_solve_mixed<0>, dim3((unsigned)batch), dim3(kLuBlock),
                           (unsigned)p.total, ctx->stream, (int)n, A, rhs, f32, max_iter, tol,
                           info, iters, p.lda, (unsigned)p.mat);
    FV_LAUNCH_CHECK(ctx);
    tl.done();
    return FV_OK;
}

}  // namespace fv

using namespace fv;

extern "C" {

int fv_dense_solve_dev(fv_ctx *ctx, int64_t batch, int64_t n, double *A, double *rhs,
                       int32_t *info)
{
    if (!ctx) return FV_E_ARG;
    FV_REQUIRE(ctx, batch >= 0 && n >= 0, "negative size");
    if (batch == 0 || n == 0) return FV_OK;
    FV_REQUIRE(ctx, A && rhs, "null matrix/rhs");
    FV_REQUIRE(ctx, n <= 1024, "dense solve supports n <= 1024 (got %lld)", (long long)n);
    return dense_solve(ctx, batch, n, A, rhs, info);
}

int fv_dense_solve_host(fv_ctx *ctx, int64_t batch, int64_t n, double *A, double *rhs,
                        int32_t *info)
{
    if (!ctx) return FV_E_ARG;
    FV_REQUIRE(ctx, batch >= 0 && n >= 0, "negative size");
    if (batch == 0 || n == 0) return FV_OK;
    FV_REQUIRE(ctx, A && rhs && info, "null matrix/rhs/info");
    double *dA = upload(ctx, S_TMP0, A, (size_t)batch * n * n);
    double *db = upload(ctx, S_OUT, rhs, (size_t)batch * n);
    int32_t *dinfo = static_cast<int32_t *>(workspace(ctx, S_INFO, sizeof(int32_t) * batch));
    if (!dA || !db || !dinfo) return FV_E_HIP;
    int rc = fv_dense_solve_dev(ctx, batch, n, dA, db, dinfo);
    if (rc) return rc;
    FV_HIP_CHECK(ctx, hipMemcpyAsync(A, dA, sizeof(double) * batch * n * n, hipMemcpyDeviceToHost,
                                     ctx->stream));
    FV_HIP_CHECK(ctx, hipMemcpyAsync(rhs, db, sizeof(double) * batch * n, hipMemcpyDeviceToHost,
                                     ctx->stream));
    FV_HIP_CHECK(ctx, hipMemcpyAsync(info, dinfo, sizeof(int32_t) * batch, hipMemcpyDeviceToHost,
                                     ctx->stream));
    FV_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return FV_OK;
}

int fv_dense_solve_mixed_dev(fv_ctx *ctx, int64_t batch, int64_t n, const double *A,
                             double *rhs, int32_t max_iter, double tol, int32_t *info,
                             int32_t *iters)
{
    if (!ctx) return FV_E_ARG;
    FV_REQUIRE(ctx, batch >= 0 && n >= 0 && max_iter >= 0, "negative size");
    if (batch == 0 || n == 0) return FV_OK;
    FV_REQUIRE(ctx, A && rhs, "null matrix/rhs");
    FV_REQUIRE(ctx, n <= 1024, "dense solve supports n <= 1024 (got %lld)", (long long)n);
    return dense_solve_mixed(ctx, batch, n, A, rhs, max_iter, tol, info, iters);
}

int fv_dense_solve_mixed_host(fv_ctx *ctx, int64_t batch, int64_t n, const double *A,
                              double *rhs, int32_t max_iter, double tol, int32_t *info,
                              int32_t *iters)
{
    if (!ctx) return FV_E_ARG;
    FV_REQUIRE(ctx, batch >= 0 && n >= 0 && max_iter >= 0, "negative size");
    if (batch == 0 || n == 0) return FV_OK;
    FV_REQUIRE(ctx, A && rhs, "null matrix/rhs");
    double *dA = upload(ctx, S_TMP0, A, (size_t)batch * n * n);
    double *db = upload(ctx, S_OUT, rhs, (size_t)batch * n);
    int32_t *dinfo = static_cast<int32_t *>(workspace(ctx, S_INFO, sizeof(int32_t) * 2 * batch));
    if (!dA || !db || !dinfo) return FV_E_HIP;
    int rc = fv_dense_solve_mixed_dev(ctx, batch, n, dA, db, max_iter, tol, dinfo, dinfo + batch);
    if (rc) return rc;
    FV_HIP_CHECK(ctx, hipMemcpyAsync(rhs, db, sizeof(double) * batch * n, hipMemcpyDeviceToHost,
                                     ctx->stream));
    if (info)
        FV_HIP_CHECK(ctx, hipMemcpyAsync(info, dinfo, sizeof(int32_t) * batch,
                                         hipMemcpyDeviceToHost, ctx->stream));
    if (iters)
        FV_HIP_CHECK(ctx, hipMemcpyAsync(iters, dinfo + batch, sizeof(int32_t) * batch,
                                         hipMemcpyDeviceToHost, ctx->stream));
    FV_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return FV_OK;
}

}  // extern "C"
