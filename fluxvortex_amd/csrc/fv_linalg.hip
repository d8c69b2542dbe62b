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
//   1. wave 0 factors the n-k0 x kb panel with partial pivoting, swapping
//      whole rows and eliminating the right-hand side as it goes (wave-
//      synchronous: no workgroup barrier inside the panel);
//   2. all waves form U12 = L11^{-1} A12 (one column per lane);
//   3. all waves update A22 -= L21 U12.
// Three workgroup barriers per panel instead of three per column.
constexpr int kNb = 16;

template <typename T>
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <typename T, typename TB>
__device__ int lu_blocked(T *a, int n, int *perm, TB *b)
{
    __shared__ int info_s;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    if (threadIdx.x == 0) info_s = 0;
    __syncthreads();
    for (int k0 = 0; k0 < n; k0 += kNb) {
        const int kb = (n - k0 < kNb) ? n - k0 : kNb;
        if (wave == 0) {
            for (int j = k0; j < k0 + kb; ++j) {
                double best = -1.0;
                int bi = n;
                for (int r = j + lane; r < n; r += 64) {
                    const double v = fabs((double)a[(int64_t)r * n + j]);
                    if (v > best) {
                        best = v;
                        bi = r;
                    }
                }
                for (int off = 32; off > 0; off >>= 1) {
                    const double ov = __shfl_xor(best, off, 64);
                    const int oi = __shfl_xor(bi, off, 64);
                    if (ov > best || (ov == best && oi < bi)) {
                        best = ov;
                        bi = oi;
                    }
                }
                const int piv = __builtin_amdgcn_readfirstlane(bi);
                if (perm && lane == 0) perm[j] = piv;
                if (best == 0.0) {
                    if (lane == 0 && info_s == 0) info_s = j + 1;
                    if (perm && lane == 0) perm[j] = j;
                    wave_sync<T>();
                    continue;
                }
                if (piv != j) {
                    for (int c = lane; c < n; c += 64) {
                        const T t = a[(int64_t)j * n + c];
                        a[(int64_t)j * n + c] = a[(int64_t)piv * n + c];
                        a[(int64_t)piv * n + c] = t;
                    }
                    if (b && lane == 0) {
                        const TB t = b[j];
                        b[j] = b[piv];
                        b[piv] = t;
                    }
                }
                wave_sync<T>();
                const T inv = (T)1 / a[(int64_t)j * n + j];
                const TB bj = b ? b[j] : (TB)0;
                for (int r = j + 1 + lane; r < n; r += 64) {
                    T *row = a + (int64_t)r * n;
                    const T l = row[j] * inv;
                    row[j] = l;
                    for (int c = j + 1; c < k0 + kb; ++c) row[c] -= l * a[(int64_t)j * n + c];
                    if (b) b[r] -= (TB)l * bj;
                }
                wave_sync<T>();
            }
        }
        __syncthreads();
        // U12 = L11^{-1} A12: forward substitution per column
        for (int c = k0 + kb + threadIdx.x; c < n; c += blockDim.x)
            for (int i = 1; i < kb; ++i) {
                T s = a[(int64_t)(k0 + i) * n + c];
                for (int m = 0; m < i; ++m)
                    s -= a[(int64_t)(k0 + i) * n + k0 + m] * a[(int64_t)(k0 + m) * n + c];
                a[(int64_t)(k0 + i) * n + c] = s;
            }
        __syncthreads();
        // A22 -= L21 U12
        const int m = n - k0 - kb;
        if (m > 0) {
            for (int e = threadIdx.x; e < m * m; e += blockDim.x) {
                const int r = k0 + kb + e / m, c = k0 + kb + e % m;
                T s = a[(int64_t)r * n + c];
                const T *lr = a + (int64_t)r * n + k0;
                for (int q = 0; q < kb; ++q) s -= lr[q] * a[(int64_t)(k0 + q) * n + c];
                a[(int64_t)r * n + c] = s;
            }
        }
        __syncthreads();
    }
    return info_s;
}

// x = U^{-1} x in place by wave 0 (row dot products, shuffle reductions).
template <typename T, typename TB>
__device__ void wave_back_subst(const T *a, int n, TB *x)
{
    if ((threadIdx.x >> 6) != 0) return;
    const int lane = threadIdx.x & 63;
    for (int k = n - 1; k >= 0; --k) {
        TB s = 0;
        for (int c = k + 1 + lane; c < n; c += 64) s += (TB)a[(int64_t)k * n + c] * x[c];
        for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
        if (lane == 0) x[k] = (x[k] - s) / (TB)a[(int64_t)k * n + k];
        wave_sync<T>();
    }
}

// Forward substitution L y = P x with the recorded pivots, by wave 0.
template <typename T, typename TB>
__device__ void wave_fwd_subst(const T *a, const int *perm, int n, TB *x)
{
    if ((threadIdx.x >> 6) != 0) return;
    const int lane = threadIdx.x & 63;
    if (lane == 0)
        for (int k = 0; k < n; ++k) {
            const int p = perm[k];
            if (p != k) {
                const TB t = x[k];
                x[k] = x[p];
                x[p] = t;
            }
        }
    wave_sync<T>();
    for (int k = 1; k < n; ++k) {
        TB s = 0;
        for (int c = lane; c < k; c += 64) s += (TB)a[(int64_t)k * n + c] * x[c];
        for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
        if (lane == 0) x[k] -= s;
        wave_sync<T>();
    }
}

__global__ __launch_bounds__(kLuBlock) void k_lu_solve_f64(int n, double *__restrict__ A,
                                                           double *__restrict__ B,
                                                           int32_t *__restrict__ info,
                                                           int use_lds)
{
    extern __shared__ double lds[];
    double *a = A + (int64_t)blockIdx.x * n * n;
    double *b = B + (int64_t)blockIdx.x * n;
    double *w = a;
    if (use_lds) {
        w = lds;
        for (int64_t e = threadIdx.x; e < (int64_t)n * n; e += kLuBlock) w[e] = a[e];
        __syncthreads();
    }
    const int inf = lu_blocked<double, double>(w, n, nullptr, b);
    if (inf == 0) wave_back_subst<double, double>(w, n, b);
    __syncthreads();
    if (use_lds)
        for (int64_t e = threadIdx.x; e < (int64_t)n * n; e += kLuBlock) a[e] = w[e];
    if (threadIdx.x == 0 && info) info[blockIdx.x] = inf;
}

// Mixed: FP32 factors in LDS (or in the global scratch f32), FP64 refinement.
__global__ __launch_bounds__(kLuBlock) void k_lu_solve_mixed(
    int n, const double *__restrict__ A, double *__restrict__ B, float *__restrict__ f32,
    double *__restrict__ work, int *__restrict__ permg, int max_iter, double tol,
    int32_t *__restrict__ info, int32_t *__restrict__ iters, int use_lds)
{
    extern __shared__ float ldsf[];
    const double *a = A + (int64_t)blockIdx.x * n * n;
    double *b = B + (int64_t)blockIdx.x * n;
    float *lu = use_lds ? ldsf : f32 + (int64_t)blockIdx.x * n * n;
    double *x = work + (int64_t)blockIdx.x * 2 * n;   // x, r
    double *r = x + n;
    int *perm = permg + (int64_t)blockIdx.x * n;
    __shared__ double stat[2];
    for (int64_t e = threadIdx.x; e < (int64_t)n * n; e += kLuBlock) lu[e] = (float)a[e];
    __syncthreads();
    const int inf = lu_blocked<float, double>(lu, n, perm, (double *)nullptr);
    if (inf != 0) {
        if (threadIdx.x == 0) {
            if (info) info[blockIdx.x] = inf;
            if (iters) iters[blockIdx.x] = 0;
        }
        return;
    }
    for (int i = threadIdx.x; i < n; i += kLuBlock) {
        x[i] = 0.0;
        r[i] = b[i];
    }
    __syncthreads();
    int it = 0;
    for (; it <= max_iter; ++it) {
        // correction d = LU32^{-1} r (carried in FP64 storage)
        wave_fwd_subst<float, double>(lu, perm, n, r);
        wave_back_subst<float, double>(lu, n, r);
        __syncthreads();
        double dmax = 0.0, xmax = 0.0;
        if (threadIdx.x < 64) {
            for (int i = threadIdx.x; i < n; i += 64) {
                x[i] += r[i];
                dmax = fmax(dmax, fabs(r[i]));
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
        // FP64 residual r = b - A x
        for (int i = threadIdx.x; i < n; i += kLuBlock) {
            double s = b[i];
            const double *row = a + (int64_t)i * n;
            for (int c = 0; c < n; ++c) s = fma(-row[c], x[c], s);
            r[i] = s;
        }
        __syncthreads();
    }
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += kLuBlock) b[i] = x[i];
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
    const size_t need = (size_t)n * n * sizeof(double) + 1024;
    const int use_lds = need <= max_lds(ctx) ? 1 : 0;
    TimedLaunch tl(ctx, "dense_solve");
    hipLaunchKernelGGL(k_lu_solve_f64, dim3((unsigned)batch), dim3(kLuBlock),
                       use_lds ? (unsigned)(n * n * sizeof(double)) : 0u, ctx->stream, (int)n, A,
                       rhs, info, use_lds);
    FV_LAUNCH_CHECK(ctx);
    tl.done();
    return FV_OK;
}

int dense_solve_mixed(fv_ctx *ctx, int64_t batch, int64_t n, const double *A, double *rhs,
                      int max_iter, double tol, int32_t *info, int32_t *iters)
{
    if (batch <= 0 || n <= 0) return FV_OK;
    const size_t need = (size_t)n * n * sizeof(float) + 1024;
    const int use_lds = need <= max_lds(ctx) ? 1 : 0;
    float *f32 = nullptr;
    if (!use_lds) {
        f32 = static_cast<float *>(
            workspace(ctx, S_MAT32, sizeof(float) * (size_t)batch * n * n));
        if (!f32) return FV_E_HIP;
    }
    double *work = static_cast<double *>(workspace(ctx, S_VEC, sizeof(double) * 2 * batch * n));
    int *perm = static_cast<int *>(workspace(ctx, S_TMP1, sizeof(int) * batch * n));
    if (!work || !perm) return FV_E_HIP;
    TimedLaunch tl(ctx, "dense_solve");
    hipLaunchKernelGGL(k_lu_solve_mixed, dim3((unsigned)batch), dim3(kLuBlock),
                       use_lds ? (unsigned)(n * n * sizeof(float)) : 0u, ctx->stream, (int)n, A,
                       rhs, f32, work, perm, max_iter, tol, info, iters, use_lds);
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
    FV_REQUIRE(ctx, n <= 4096, "dense solve supports n <= 4096 (got %lld)", (long long)n);
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
    FV_REQUIRE(ctx, n <= 4096, "dense solve supports n <= 4096 (got %lld)", (long long)n);
    return dense_solve_mixed(ctx, batch, n, A, rhs, max_iter, tol, info, iters);
}

}  // extern "C"
