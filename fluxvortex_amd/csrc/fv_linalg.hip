// fv_linalg.hip -- batched dense circulation solves.
//
// One workgroup per system (the reference's batched solver runs one thread
// per system, warp_fsi/batched_solver.py:20-63).  LU with partial pivoting
// (first maximum wins, as LAPACK's idamax), the matrix held in LDS when it
// fits and in global memory (L2-resident at these sizes) otherwise.
//   FP64:  factor + solve in FP64 (np.linalg.solve, standalone_uvlm.py:427).
//   Mixed: factor in FP32, then iterative refinement with FP64 residuals
//          r = b - A x against the FP64 matrix.
#include <cmath>

#include "fv_launch.h"

namespace fv {

constexpr int kLuBlock = 256;

template <typename T>
__device__ void block_argmax(const T *a, int n, int k, int *piv_out, T *best_out)
{
    __shared__ double sval[kLuBlock / 64];
    __shared__ int sidx[kLuBlock / 64];
    double best = -1.0;
    int bi = n;
    for (int r = k + threadIdx.x; r < n; r += kLuBlock) {
        const double v = fabs((double)a[(int64_t)r * n + k]);
        if (v > best) {
            best = v;
            bi = r;
        }
    }
    // wave reduction (first maximum wins)
    for (int off = 32; off > 0; off >>= 1) {
        const double ov = __shfl_xor(best, off, 64);
        const int oi = __shfl_xor(bi, off, 64);
        if (ov > best || (ov == best && oi < bi)) {
            best = ov;
            bi = oi;
        }
    }
    const int w = threadIdx.x / 64;
    if ((threadIdx.x & 63) == 0) {
        sval[w] = best;
        sidx[w] = bi;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double b = sval[0];
        int i = sidx[0];
        for (int q = 1; q < kLuBlock / 64; ++q) {
            if (sval[q] > b || (sval[q] == b && sidx[q] < i)) {
                b = sval[q];
                i = sidx[q];
            }
        }
        *piv_out = i;
        *best_out = (T)b;
    }
    __syncthreads();
}

// In-place LU with partial pivoting of the n x n row-major matrix a, applied
// to the right-hand side b as it goes.  Returns info (0 or k+1).
template <typename T, typename TB>
__device__ int lu_factor_apply(T *a, int n, int *perm, TB *b)
{
    __shared__ int piv_s;
    __shared__ T best_s;
    int info = 0;
    for (int k = 0; k < n; ++k) {
        block_argmax<T>(a, n, k, &piv_s, &best_s);
        const int piv = piv_s;
        if (best_s == (T)0) {
            if (info == 0) info = k + 1;
            if (threadIdx.x == 0 && perm) perm[k] = k;
            __syncthreads();
            continue;
        }
        if (threadIdx.x == 0 && perm) perm[k] = piv;
        if (piv != k) {
            for (int c = threadIdx.x; c < n; c += kLuBlock) {
                const T t = a[(int64_t)k * n + c];
                a[(int64_t)k * n + c] = a[(int64_t)piv * n + c];
                a[(int64_t)piv * n + c] = t;
            }
            if (b && threadIdx.x == 0) {
                const TB t = b[k];
                b[k] = b[piv];
                b[piv] = t;
            }
        }
        __syncthreads();
        const T inv = (T)1 / a[(int64_t)k * n + k];
        for (int r = k + 1 + threadIdx.x; r < n; r += kLuBlock) a[(int64_t)r * n + k] *= inv;
        __syncthreads();
        const int m = n - k - 1;
        for (int64_t e = threadIdx.x; e < (int64_t)m * m; e += kLuBlock) {
            const int r = k + 1 + (int)(e / m), c = k + 1 + (int)(e % m);
            a[(int64_t)r * n + c] -= a[(int64_t)r * n + k] * a[(int64_t)k * n + c];
        }
        if (b)
            for (int r = k + 1 + threadIdx.x; r < n; r += kLuBlock)
                b[r] -= (TB)a[(int64_t)r * n + k] * b[k];
        __syncthreads();
    }
    return info;
}

// Back substitution U x = b (b overwritten), column-oriented.
template <typename T, typename TB>
__device__ void lu_back(const T *a, int n, TB *b)
{
    for (int k = n - 1; k >= 0; --k) {
        __syncthreads();
        if (threadIdx.x == 0) b[k] = b[k] / (TB)a[(int64_t)k * n + k];
        __syncthreads();
        const TB xk = b[k];
        for (int r = threadIdx.x; r < k; r += kLuBlock) b[r] -= (TB)a[(int64_t)r * n + k] * xk;
    }
    __syncthreads();
}

// Forward substitution with unit-lower L and the recorded row swaps.
template <typename T, typename TB>
__device__ void lu_forward(const T *a, const int *perm, int n, TB *b)
{
    for (int k = 0; k < n; ++k) {
        __syncthreads();
        if (threadIdx.x == 0) {
            const int p = perm[k];
            if (p != k) {
                const TB t = b[k];
                b[k] = b[p];
                b[p] = t;
            }
        }
        __syncthreads();
        const TB bk = b[k];
        for (int r = k + 1 + threadIdx.x; r < n; r += kLuBlock)
            b[r] -= (TB)a[(int64_t)r * n + k] * bk;
    }
    __syncthreads();
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
    const int inf = lu_factor_apply<double, double>(w, n, nullptr, b);
    if (inf == 0) lu_back<double, double>(w, n, b);
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
    const int inf = lu_factor_apply<float, double>(lu, n, perm, (double *)nullptr);
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
    int it = 0;
    for (; it <= max_iter; ++it) {
        __syncthreads();
        // correction d = LU32^{-1} r (carried in FP64 storage)
        lu_forward<float, double>(lu, perm, n, r);
        lu_back<float, double>(lu, n, r);
        if (threadIdx.x == 0) {
            stat[0] = 0.0;
            stat[1] = 0.0;
        }
        __syncthreads();
        double dmax = 0.0, xmax = 0.0;
        for (int i = threadIdx.x; i < n; i += kLuBlock) {
            x[i] += r[i];
            dmax = fmax(dmax, fabs(r[i]));
            xmax = fmax(xmax, fabs(x[i]));
        }
        for (int off = 32; off > 0; off >>= 1) {
            dmax = fmax(dmax, __shfl_xor(dmax, off, 64));
            xmax = fmax(xmax, __shfl_xor(xmax, off, 64));
        }
        if ((threadIdx.x & 63) == 0) {
            atomicMax((unsigned long long *)&stat[0], (unsigned long long)__double_as_longlong(dmax));
            atomicMax((unsigned long long *)&stat[1], (unsigned long long)__double_as_longlong(xmax));
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
