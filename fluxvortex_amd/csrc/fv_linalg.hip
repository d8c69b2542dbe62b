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

#include "fv_lu.h"

namespace fv {

constexpr int kLuBlock = 256;

// LDS plan of one solve: the whole matrix (odd row stride) when it fits,
// else a staged panel; then the vectors.
struct LuLds {
    int in_lds;      // matrix in LDS
    int lda;         // row stride of the matrix
    size_t mat;      // bytes of the matrix or panel area
    size_t total;    // dynamic LDS bytes
};

__host__ __device__ inline size_t lu_align(size_t v) { return (v + 15) & ~size_t(15); }

// vec_lds / vec_glob: bytes of the vectors when the matrix is in LDS / in
// global memory; limit: the dynamic LDS a workgroup may take (the kernels'
// own static LDS already subtracted).
template <typename T>
static LuLds lu_lds_plan(int64_t n, size_t vec_lds, size_t vec_glob, size_t limit)
{
    LuLds p;
    const int64_t ldo = n | 1;
    const size_t full = lu_align((size_t)n * ldo * sizeof(T));
    if (full + vec_lds <= limit) {
        p.in_lds = 1;
        p.lda = (int)ldo;
        p.mat = full;
        p.total = p.mat + vec_lds;
    } else {
        p.in_lds = 0;
        p.lda = (int)n;
        p.mat = lu_align((size_t)n * kLdp * sizeof(T));
        p.total = p.mat + vec_glob;
    }
    return p;
}

// IN_LDS: the matrix is copied into LDS (row stride lda_lds) and factored
// there; else it is factored in global memory with the panels staged in LDS.
// (A compile-time choice, so LDS accesses compile to ds_* instructions.)
template <int IN_LDS>
__global__ __launch_bounds__(kLuBlock) void k_lu_solve_f64(int n, double *__restrict__ A,
                                                           double *__restrict__ B,
                                                           int32_t *__restrict__ info,
                                                           int lda_lds, unsigned mat_bytes,
                                                           LuReuse ru, int store_factors)
{
    extern __shared__ __attribute__((aligned(16))) unsigned char lds_raw[];
    const int64_t nn = (int64_t)n * n;
    double *a = A + (int64_t)blockIdx.x * nn;
    double *bg = B + (int64_t)blockIdx.x * n;
    double *area = reinterpret_cast<double *>(lds_raw);
    double *b = reinterpret_cast<double *>(lds_raw + mat_bytes);
    double *rdiag = b + n;
    int *perm = reinterpret_cast<int *>(rdiag + n);
    // the row order: after the pivots in LDS, or in the staged-panel area
    // (free once the factorization is done) when the matrix is in global memory
    int *order = IN_LDS ? perm + n : reinterpret_cast<int *>(area);
    double *w = IN_LDS ? area : a;
    const int lda = IN_LDS ? lda_lds : n;
    double *key = ru.key ? ru.key + (int64_t)blockIdx.x * nn : nullptr;
    double *fac = ru.key ? static_cast<double *>(ru.fac) + (int64_t)blockIdx.x * nn : nullptr;
    int32_t *ord = ru.key ? ru.order + (int64_t)blockIdx.x * n : nullptr;
    const bool check = key && ru.state[blockIdx.x] == 1;
    const bool differ = mat_move<kLuBlock>(a, n, IN_LDS ? w : (double *)nullptr, lda, n,
                                           check ? key : (const double *)nullptr);
    const bool reuse = !__syncthreads_or((!check || differ) ? 1 : 0);
    int inf = 0;
    if (reuse) {
        // the stored factors and row order of this matrix
        mat_move<kLuBlock>(fac, n, w, lda, n, (const double *)nullptr);
        for (int i = threadIdx.x; i < n; i += kLuBlock) order[i] = ord[i];
        if (threadIdx.x == 0 && ru.hits) ru.hits[blockIdx.x] += 1;
    } else {
        if (key) {
            mat_move<kLuBlock>(a, n, key, n, n, (const double *)nullptr);
            if (threadIdx.x == 0) ru.state[blockIdx.x] = 0;
        }
        inf = lu_blocked<double, double>(w, lda, n, perm, (double *)nullptr,
                                         IN_LDS ? nullptr : area);
        // the pivot sequence composed into one row order: (P v)[i] = v[order[i]]
        for (int i = threadIdx.x; i < n; i += kLuBlock) order[pivot_track(perm, n, i)] = i;
        if (key && inf == 0) {
            mat_move<kLuBlock>(w, lda, fac, n, n, (const double *)nullptr);
            __syncthreads();
            for (int i = threadIdx.x; i < n; i += kLuBlock) ord[i] = order[i];
            if (threadIdx.x == 0) ru.state[blockIdx.x] = 1;
        }
    }
    __syncthreads();
    // both paths solve with the same two substitutions: L y = P b, U x = y
    for (int i = threadIdx.x; i < n; i += kLuBlock) {
        b[i] = bg[order[i]];
        rdiag[i] = 1.0 / w[(int64_t)i * lda + i];
    }
    __syncthreads();
    if (inf == 0) {
        wave_fwd_subst<double, double>(w, lda, n, b);
        wave_back_subst<double, double>(w, lda, n, b, rdiag);
    }
    __syncthreads();
    // A takes its factors (LAPACK's getrf / gesv contract); in global memory
    // it holds the stored factors after a reuse
    if (IN_LDS && store_factors) mat_move<kLuBlock>(w, lda, a, n, n, (const double *)nullptr);
    for (int i = threadIdx.x; i < n; i += kLuBlock) bg[i] = b[i];
    if (threadIdx.x == 0 && info) info[blockIdx.x] = inf;
}

// n <= 128: the same solve by NW wavefronts with the matrix factored in
// registers (lu_regs); the matrix is staged through LDS (coalesced loads and
// the bitwise comparison with the stored matrix on the way in), the factors
// come back to LDS in LAPACK's row order, and the substitutions and the
// stored factors are those of the kernel above.
template <int NW>
__global__ __launch_bounds__(64 * NW) void k_lu_solve_reg(int n, double *__restrict__ A,
                                                          double *__restrict__ B,
                                                          int32_t *__restrict__ info, int lda,
                                                          unsigned mat_bytes, LuReuse ru,
                                                          int store_factors)
{
    constexpr int kLuRegLanes = 64 * NW;
    extern __shared__ __attribute__((aligned(16))) unsigned char lds_raw[];
    const int64_t nn = (int64_t)n * n;
    double *a = A + (int64_t)blockIdx.x * nn;
    double *bg = B + (int64_t)blockIdx.x * n;
    double *w = reinterpret_cast<double *>(lds_raw);
    double *b = reinterpret_cast<double *>(lds_raw + mat_bytes);
    double *rdiag = b + kLuRegMax;
    int *order = reinterpret_cast<int *>(rdiag + kLuRegMax);
    auto *sh = reinterpret_cast<LuRegShared<double> *>(order + kLuRegMax);
    double *key = ru.key ? ru.key + (int64_t)blockIdx.x * nn : nullptr;
    double *fac = ru.key ? static_cast<double *>(ru.fac) + (int64_t)blockIdx.x * nn : nullptr;
    int32_t *ord = ru.key ? ru.order + (int64_t)blockIdx.x * n : nullptr;
    const bool check = key && ru.state[blockIdx.x] == 1;
    const bool differ =
        mat_move<kLuRegLanes>(a, n, w, lda, n, check ? key : (const double *)nullptr);
    const bool reuse = !__syncthreads_or((!check || differ) ? 1 : 0);
    int inf = 0;
    if (reuse) {
        mat_move<kLuRegLanes>(fac, n, w, lda, n, (const double *)nullptr);
        for (int i = threadIdx.x; i < n; i += kLuRegLanes) order[i] = ord[i];
        if (threadIdx.x == 0 && ru.hits) ru.hits[blockIdx.x] += 1;
    } else {
        if (key) {
            mat_move<kLuRegLanes>(w, lda, key, n, n, (const double *)nullptr);
            if (threadIdx.x == 0) ru.state[blockIdx.x] = 0;
        }
        inf = lu_regs<NW, double>(w, lda, n, order, sh);
        if (key && inf == 0) {
            mat_move<kLuRegLanes>(w, lda, fac, n, n, (const double *)nullptr);
            for (int i = threadIdx.x; i < n; i += kLuRegLanes) ord[i] = order[i];
            if (threadIdx.x == 0) ru.state[blockIdx.x] = 1;
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += kLuRegLanes) {
        b[i] = bg[order[i]];
        rdiag[i] = 1.0 / w[i * lda + i];
    }
    __syncthreads();
    if (inf == 0) {
        wave_fwd_subst128<double, double>(w, lda, n, b);
        wave_back_subst128<double, double>(w, lda, n, b, rdiag);
    }
    __syncthreads();
    if (store_factors) mat_move<kLuRegLanes>(w, lda, a, n, n, (const double *)nullptr);
    for (int i = threadIdx.x; i < n; i += kLuRegLanes) bg[i] = b[i];
    if (threadIdx.x == 0 && info) info[blockIdx.x] = inf;
}

// Mixed: FP32 factors (LDS or the global scratch f32), FP64 refinement with
// the vectors in LDS.
template <int IN_LDS>
__global__ __launch_bounds__(kLuBlock) void k_lu_solve_mixed(
    int n, const double *__restrict__ A, double *__restrict__ B, float *__restrict__ f32,
    int max_iter, double tol, int32_t *__restrict__ info, int32_t *__restrict__ iters,
    int lda_lds, unsigned mat_bytes, LuReuse ru)
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
    const int64_t nn = (int64_t)n * n;
    double *key = ru.key ? ru.key + (int64_t)blockIdx.x * nn : nullptr;
    float *fac = ru.key ? static_cast<float *>(ru.fac) + (int64_t)blockIdx.x * nn : nullptr;
    int32_t *ord = ru.key ? ru.order + (int64_t)blockIdx.x * n : nullptr;
    const bool check = key && ru.state[blockIdx.x] == 1;
    const bool differ = mat_move<kLuBlock>(a, n, lu, lda, n, check ? key : (const double *)nullptr);
    const bool reuse = !__syncthreads_or((!check || differ) ? 1 : 0);
    if (reuse) {
        // the stored FP32 factors and row order of this matrix
        mat_move<kLuBlock>(fac, n, lu, lda, n, (const double *)nullptr);
        for (int i = threadIdx.x; i < n; i += kLuBlock) order[i] = ord[i];
        if (threadIdx.x == 0 && ru.hits) ru.hits[blockIdx.x] += 1;
    } else {
        if (key) {
            mat_move<kLuBlock>(a, n, key, n, n, (const double *)nullptr);
            if (threadIdx.x == 0) ru.state[blockIdx.x] = 0;
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
        // the pivot sequence composed into one row order: (P v)[i] = v[order[i]]
        for (int i = threadIdx.x; i < n; i += kLuBlock) order[pivot_track(perm, n, i)] = i;
        if (key) {
            mat_move<kLuBlock>(lu, lda, fac, n, n, (const double *)nullptr);
            __syncthreads();
            for (int i = threadIdx.x; i < n; i += kLuBlock) ord[i] = order[i];
            if (threadIdx.x == 0) ru.state[blockIdx.x] = 1;
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += kLuBlock) {
        x[i] = 0.0;
        r[i] = bg[i];
        rdiag[i] = 1.0 / (double)lu[(int64_t)i * lda + i];
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
        // FP64 residual r = b - A x: one row per lane, the row's loads
        // independent of each other (8 in flight), x broadcast from LDS;
        // rows beyond the workgroup's lanes take further passes
        for (int i = threadIdx.x; i < n; i += kLuBlock) {
            const double *row = a + (int64_t)i * n;
            double s0 = 0.0, s1 = 0.0;
            int c = 0;
            for (; c + 8 <= n; c += 8) {
                double v[8];
#pragma unroll
                for (int k = 0; k < 8; ++k) v[k] = row[c + k];
#pragma unroll
                for (int k = 0; k < 8; k += 2) {
                    s0 = fma(v[k], x[c + k], s0);
                    s1 = fma(v[k + 1], x[c + k + 1], s1);
                }
            }
            for (; c < n; ++c) s0 = fma(row[c], x[c], s0);
            r[i] = bg[i] - (s0 + s1);
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

// Mixed, n <= 128: FP32 factors from the register factorization (lu_regs),
// the substitutions in FP64 on them, and the FP64 residual r = b - A x from
// the matrix held in the same register layout (lane: rows l, l + 64;
// wavefront: columns w + NW mj), the NW wavefronts' partial row sums added
// in LDS in wavefront order.  Refinement as in the kernel above.
template <int NW>
__global__ __launch_bounds__(64 * NW) void k_lu_solve_mixed_reg(
    int n, const double *__restrict__ A, double *__restrict__ B, int max_iter, double tol,
    int32_t *__restrict__ info, int32_t *__restrict__ iters, int lda, unsigned mat_bytes,
    LuReuse ru)
{
    constexpr int kLuRegLanes = 64 * NW;
    constexpr int CW = kLuRegMax / NW;
    extern __shared__ __attribute__((aligned(16))) unsigned char lds_raw[];
    const int lane = threadIdx.x & 63;
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int64_t nn = (int64_t)n * n;
    const double *a = A + (int64_t)blockIdx.x * nn;
    double *bg = B + (int64_t)blockIdx.x * n;
    float *lu = reinterpret_cast<float *>(lds_raw);
    double *x = reinterpret_cast<double *>(lds_raw + mat_bytes);
    double *r = x + kLuRegMax;
    double *d = r + kLuRegMax;
    double *rdiag = d + kLuRegMax;
    double *part = rdiag + kLuRegMax;                        // [NW][128] partial row sums
    int *order = reinterpret_cast<int *>(part + NW * kLuRegMax);
    auto *sh = reinterpret_cast<LuRegShared<float> *>(order + kLuRegMax);
    __shared__ double stat[2];
    double *key = ru.key ? ru.key + (int64_t)blockIdx.x * nn : nullptr;
    float *fac = ru.key ? static_cast<float *>(ru.fac) + (int64_t)blockIdx.x * nn : nullptr;
    int32_t *ord = ru.key ? ru.order + (int64_t)blockIdx.x * n : nullptr;
    const bool check = key && ru.state[blockIdx.x] == 1;
    const bool differ =
        mat_move<kLuRegLanes>(a, n, lu, lda, n, check ? key : (const double *)nullptr);
    const bool reuse = !__syncthreads_or((!check || differ) ? 1 : 0);
    if (reuse) {
        mat_move<kLuRegLanes>(fac, n, lu, lda, n, (const double *)nullptr);
        for (int i = threadIdx.x; i < n; i += kLuRegLanes) order[i] = ord[i];
        if (threadIdx.x == 0 && ru.hits) ru.hits[blockIdx.x] += 1;
    } else {
        if (key) {
            mat_move<kLuRegLanes>(a, n, key, n, n, (const double *)nullptr);
            if (threadIdx.x == 0) ru.state[blockIdx.x] = 0;
        }
        const int inf = lu_regs<NW, float>(lu, lda, n, order, sh);
        if (inf != 0) {
            if (threadIdx.x == 0) {
                if (info) info[blockIdx.x] = inf;
                if (iters) iters[blockIdx.x] = 0;
            }
            return;
        }
        if (key) {
            mat_move<kLuRegLanes>(lu, lda, fac, n, n, (const double *)nullptr);
            for (int i = threadIdx.x; i < n; i += kLuRegLanes) ord[i] = order[i];
            if (threadIdx.x == 0) ru.state[blockIdx.x] = 1;
        }
    }
    // the FP64 matrix in the register layout, for the residuals
    double av[2][CW];
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        const int row = lane + 64 * q;
#pragma unroll
        for (int mj = 0; mj < CW; ++mj) {
            const int j = wv + NW * mj;
            av[q][mj] = (row < n && j < n) ? a[row * n + j] : 0.0;
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += kLuRegLanes) {
        x[i] = 0.0;
        r[i] = bg[i];
        rdiag[i] = 1.0 / (double)lu[i * lda + i];
    }
    __syncthreads();
    int it = 0;
    for (; it <= max_iter; ++it) {
        // correction d = U32^{-1} L32^{-1} P r (carried in FP64 storage)
        for (int i = threadIdx.x; i < n; i += kLuRegLanes) d[i] = r[order[i]];
        __syncthreads();
        wave_fwd_subst128<float, double>(lu, lda, n, d);
        wave_back_subst128<float, double>(lu, lda, n, d, rdiag);
        __syncthreads();
        if (wv == 0) {
            double dmax = 0.0, xmax = 0.0;
            for (int i = lane; i < n; i += 64) {
                x[i] += d[i];
                dmax = fmax(dmax, fabs(d[i]));
                xmax = fmax(xmax, fabs(x[i]));
            }
            for (int off = 32; off > 0; off >>= 1) {
                dmax = fmax(dmax, __shfl_xor(dmax, off, 64));
                xmax = fmax(xmax, __shfl_xor(xmax, off, 64));
            }
            if (lane == 0) {
                stat[0] = dmax;
                stat[1] = xmax;
            }
        }
        __syncthreads();
        const bool done = stat[0] <= tol * stat[1];
        if (done || it == max_iter) break;
        // FP64 residual r = b - A x
        double p0 = 0.0, p1 = 0.0;
#pragma unroll
        for (int mj = 0; mj < CW; ++mj) {
            const int j = wv + NW * mj;
            const double xj = j < n ? x[j] : 0.0;
            p0 = fma(av[0][mj], xj, p0);
            p1 = fma(av[1][mj], xj, p1);
        }
        part[wv * kLuRegMax + lane] = p0;
        part[wv * kLuRegMax + lane + 64] = p1;
        __syncthreads();
        for (int i = threadIdx.x; i < n; i += kLuRegLanes) {
            double s = 0.0;
#pragma unroll
            for (int w = 0; w < NW; ++w) s += part[w * kLuRegMax + i];
            r[i] = bg[i] - s;
        }
        __syncthreads();
    }
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += kLuRegLanes) bg[i] = x[i];
    if (threadIdx.x == 0) {
        if (info) info[blockIdx.x] = 0;
        if (iters) iters[blockIdx.x] = it;
    }
}

// Dynamic LDS a solve may take: the device's per-workgroup limit less a
// margin for the kernels' static LDS (a few hundred bytes).
constexpr size_t kLuStaticLds = 1024;

static size_t max_lds(fv_ctx *ctx)
{
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, ctx->device) !=
            hipSuccess ||
        v <= 0)
        v = 65536;
    return (size_t)v - kLuStaticLds;
}

// Wavefronts of the register solve (n <= 128).
constexpr int kLuRegWaves = 16;

// store_factors: A takes its LU factors (else A is left as it was when the
// matrix is factored in LDS).
int dense_solve(fv_ctx *ctx, int64_t batch, int64_t n, double *A, double *rhs, int32_t *info,
                const LuReuse *reuse, bool store_factors)
{
    const LuReuse ru = reuse ? *reuse : LuReuse{nullptr, nullptr, nullptr, nullptr, nullptr};
    if (batch <= 0 || n <= 0) return FV_OK;
    if (n <= kLuRegMax) {
        const int lda = (int)(n | 1);
        const size_t mat = lu_align((size_t)n * lda * sizeof(double));
        const size_t total = mat + 2 * kLuRegMax * sizeof(double) + kLuRegMax * sizeof(int) +
                             sizeof(LuRegShared<double>);
        FV_REQUIRE(ctx, total <= max_lds(ctx), "dense solve: %zu bytes of LDS for n = %lld",
                   total, (long long)n);
        TimedLaunch tl(ctx, "dense_solve");
        hipLaunchKernelGGL(k_lu_solve_reg<kLuRegWaves>, dim3((unsigned)batch),
                           dim3(64 * kLuRegWaves), (unsigned)total, ctx->stream, (int)n, A, rhs,
                           info, lda, (unsigned)mat, ru, store_factors ? 1 : 0);
        FV_LAUNCH_CHECK(ctx);
        tl.done();
        return FV_OK;
    }
    const size_t vec = 2 * lu_align((size_t)n * sizeof(double)) + lu_align((size_t)n * sizeof(int));
    const LuLds p = lu_lds_plan<double>(n, vec + lu_align((size_t)n * sizeof(int)), vec,
                                        max_lds(ctx));
    FV_REQUIRE(ctx, p.total <= max_lds(ctx), "dense solve: n = %lld exceeds the LDS panel limit",
               (long long)n);
    TimedLaunch tl(ctx, "dense_solve");
    if (p.in_lds)
        hipLaunchKernelGGL(k_lu_solve_f64<1>, dim3((unsigned)batch), dim3(kLuBlock),
                           (unsigned)p.total, ctx->stream, (int)n, A, rhs, info, p.lda,
                           (unsigned)p.mat, ru, store_factors ? 1 : 0);
    else
        hipLaunchKernelGGL(k_lu_solve_f64<0>, dim3((unsigned)batch), dim3(kLuBlock),
                           (unsigned)p.total, ctx->stream, (int)n, A, rhs, info, p.lda,
                           (unsigned)p.mat, ru, 1);
    FV_LAUNCH_CHECK(ctx);
    tl.done();
    return FV_OK;
}

int dense_solve_mixed(fv_ctx *ctx, int64_t batch, int64_t n, const double *A, double *rhs,
                      int max_iter, double tol, int32_t *info, int32_t *iters,
                      const LuReuse *reuse)
{
    const LuReuse ru = reuse ? *reuse : LuReuse{nullptr, nullptr, nullptr, nullptr, nullptr};
    if (batch <= 0 || n <= 0) return FV_OK;
    if (n <= kLuRegMax) {
        const int lda = (int)(n | 1);
        const size_t mat = lu_align((size_t)n * lda * sizeof(float));
        const size_t total = mat + (4 + kLuRegWaves) * kLuRegMax * sizeof(double) +
                             kLuRegMax * sizeof(int) + sizeof(LuRegShared<float>);
        FV_REQUIRE(ctx, total <= max_lds(ctx), "dense solve: %zu bytes of LDS for n = %lld",
                   total, (long long)n);
        TimedLaunch tl(ctx, "dense_solve");
        hipLaunchKernelGGL(k_lu_solve_mixed_reg<kLuRegWaves>, dim3((unsigned)batch),
                           dim3(64 * kLuRegWaves), (unsigned)total, ctx->stream, (int)n, A, rhs,
                           max_iter, tol, info, iters, lda, (unsigned)mat, ru);
        FV_LAUNCH_CHECK(ctx);
        tl.done();
        return FV_OK;
    }
    const size_t vec = 4 * lu_align((size_t)n * sizeof(double)) +
                       2 * lu_align((size_t)n * sizeof(int));
    const LuLds p = lu_lds_plan<float>(n, vec, vec, max_lds(ctx));
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
                           info, iters, p.lda, (unsigned)p.mat, ru);
    else
        hipLaunchKernelGGL(k_lu_solve_mixed<0>, dim3((unsigned)batch), dim3(kLuBlock),
                           (unsigned)p.total, ctx->stream, (int)n, A, rhs, f32, max_iter, tol,
                           info, iters, p.lda, (unsigned)p.mat, ru);
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
    return dense_solve(ctx, batch, n, A, rhs, info, nullptr, true);
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
    return dense_solve_mixed(ctx, batch, n, A, rhs, max_iter, tol, info, iters, nullptr);
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
