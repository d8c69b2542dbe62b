// fv_mlp.hip -- the far-field closure network: a 32 -> 128 -> 128 -> 128 -> 16
// SiLU MLP with BF16 weights/activations on MFMA (v_mfma_f32_16x16x16_bf16,
// FP32 accumulation), and its hand-written backward and Adam kernels.
//
// One wavefront owns 16 samples at a time; its activations live in LDS as
// BF16 (the A operand), weights are read as BF16 rows (the B operand), and the
// 16x16 FP32 accumulator tiles are biased, activated and rounded to BF16 in
// registers before the next layer.  Inputs are zero-padded to 32 columns and
// outputs to 16 (two and one 16-wide MFMA k / n slices).
#include <cmath>
#include <cstring>
#include <new>
#include <vector>

#include "fv_launch.h"

typedef short fv_s4 __attribute__((ext_vector_type(4)));
typedef float fv_f4 __attribute__((ext_vector_type(4)));

namespace fv {

constexpr int kIn = 32, kHid = 128, kOut = 16;
// FP32 master parameter layout (row-major weights [out][in]).
constexpr int kOffW1 = 0;
constexpr int kOffB1 = kOffW1 + kHid * kIn;
constexpr int kOffW2 = kOffB1 + kHid;
constexpr int kOffB2 = kOffW2 + kHid * kHid;
constexpr int kOffW3 = kOffB2 + kHid;
constexpr int kOffB3 = kOffW3 + kHid * kHid;
constexpr int kOffW4 = kOffB3 + kHid;
constexpr int kOffB4 = kOffW4 + kOut * kHid;
constexpr int kParams = kOffB4 + kOut;   // 39312
constexpr int kLda = kHid + 8;           // LDS row stride of activations (bf16)

__device__ __forceinline__ unsigned short f2bf(float f)
{
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);     // round to nearest even
    return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf2f(unsigned short h) { return __uint_as_float((uint32_t)h << 16); }
__device__ __forceinline__ float sigm(float z) { return 1.0f / (1.0f + expf(-z)); }
__device__ __forceinline__ float silu(float z) { return z * sigm(z); }
__device__ __forceinline__ float dsilu(float z)
{
    const float s = sigm(z);
    return s * (1.0f + z * (1.0f - s));
}

struct MlpW {
    const unsigned short *w[4];    // bf16 [out][in]
    const unsigned short *wt[4];   // bf16 [in][out] (backward)
    const float *b[4];
};

// acc[t] (16 samples x 16 outputs, tile t) = A (16 x K, LDS bf16, stride lda)
//                                           x W[t*16 .. t*16+15][0..K)^T
// MFMA 16x16x16 operand layout: lane l supplies row (l & 15) and the four
// k values 4 (l >> 4) .. +3 of the current 16-wide k slice; the result holds
// rows 4 (l >> 4) + v, column (l & 15).
template <int K, int NT>
__device__ __forceinline__ void wave_gemm_nt(const unsigned short *A, int lda,
                                             const unsigned short *__restrict__ W, int ldw,
                                             fv_f4 (&acc)[NT])
{
    const int l = threadIdx.x & 63;
    const int r = l & 15, kq = (l >> 4) * 4;
#pragma unroll
    for (int t = 0; t < NT; ++t) acc[t] = fv_f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int k0 = 0; k0 < K; k0 += 16) {
        const fv_s4 a = *reinterpret_cast<const fv_s4 *>(A + r * lda + k0 + kq);
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            const fv_s4 b = *reinterpret_cast<const fv_s4 *>(W + (t * 16 + r) * ldw + k0 + kq);
            acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[t], 0, 0, 0);
        }
    }
}

// LDS plan of the forward and backward kernels: the bf16 weight matrices of
// the pass, staged once per workgroup with padded row strides (K + 8, so the
// 16 rows a wavefront reads per MFMA step fall in distinct banks), then per
// wavefront two 16 x kLda activation buffers.  Up to kMlpWaves wavefronts
// share the staged weights, each owning 16 samples at a time.
constexpr int kMlpWaves = 8;
constexpr int kLdw1 = kIn + 8;    // W1 [128][32]
constexpr int kLdwH = kHid + 8;   // W2, W3 [128][128]; W4 [16][128]
constexpr int kLdwT = kOut + 8;   // W4^T [128][16]
// forward: W1, W2, W3, W4
constexpr int kFw1 = 0, kFw2 = kFw1 + kHid * kLdw1, kFw3 = kFw2 + kHid * kLdwH,
              kFw4 = kFw3 + kHid * kLdwH, kFwEnd = kFw4 + kOut * kLdwH;
// backward: W2^T, W3^T, W4^T
constexpr int kBw1 = 0, kBw2 = kBw1 + kHid * kLdwH, kBw3 = kBw2 + kHid * kLdwH,
              kBwEnd = kBw3 + kHid * kLdwT;
constexpr int kActPerWave = 2 * 16 * kLda;
__host__ __device__ constexpr size_t fwd_lds(int waves)
{
    return sizeof(unsigned short) * ((size_t)kFwEnd + (size_t)waves * kActPerWave);
}
__host__ __device__ constexpr size_t bwd_lds(int waves)
{
    return sizeof(unsigned short) * ((size_t)kBwEnd + (size_t)waves * kActPerWave);
}

__device__ __forceinline__ void mlp_wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// rows x K bf16 (dense) -> LDS with row stride ld, 4 values per lane and step
__device__ __forceinline__ void stage_rows(const unsigned short *__restrict__ src, int rows,
                                           int K, unsigned short *dst, int ld)
{
    const int q = K / 4;
    for (int e = threadIdx.x; e < rows * q; e += blockDim.x) {
        const int r = e / q, c = (e - r * q) * 4;
        *reinterpret_cast<fv_s4 *>(dst + r * ld + c) =
            *reinterpret_cast<const fv_s4 *>(src + r * K + c);
    }
}

// Forward pass of nb samples; each wavefront takes 16-sample groups in turn.
//   X (nb x 16 fp32) -> Y (nb x 16 fp32).  When zs / hs are non-null the
//   pre-activations (fp32) and activations (bf16) of the 3 hidden layers are
//   kept for the backward pass: zs[(layer * nb + s) * 128 + n].
// STAGE: the weights are staged in LDS (large batches); else every wavefront
// reads them from global memory (L2), which small batches pay less for.
template <bool STAGE>
__global__ __launch_bounds__(STAGE ? 64 * kMlpWaves : 64) void k_mlp_forward(MlpW m,
                                                                const float *__restrict__ X,
                                                                int64_t nb, float *__restrict__ Y,
                                                                float *__restrict__ zs,
                                                                unsigned short *__restrict__ hs)
{
    extern __shared__ __attribute__((aligned(16))) unsigned short smem[];
    const unsigned short *wl[4] = {m.w[0], m.w[1], m.w[2], m.w[3]};
    const int ld1 = STAGE ? kLdw1 : kIn, ldh = STAGE ? kLdwH : kHid;
    if (STAGE) {
        wl[0] = smem + kFw1;
        wl[1] = smem + kFw2;
        wl[2] = smem + kFw3;
        wl[3] = smem + kFw4;
        stage_rows(m.w[0], kHid, kIn, smem + kFw1, kLdw1);
        stage_rows(m.w[1], kHid, kHid, smem + kFw2, kLdwH);
        stage_rows(m.w[2], kHid, kHid, smem + kFw3, kLdwH);
        stage_rows(m.w[3], kOut, kHid, smem + kFw4, kLdwH);
        __syncthreads();
    }
    const int wv = threadIdx.x >> 6;
    unsigned short(*act)[16 * kLda] = reinterpret_cast<unsigned short(*)[16 * kLda]>(
        smem + (STAGE ? kFwEnd : 0) + wv * kActPerWave);
    const int l = threadIdx.x & 63;
    const int col = l & 15, rq = (l >> 4) * 4;
    const int nw = blockDim.x >> 6;
    for (int64_t g = (int64_t)blockIdx.x * nw + wv; g * 16 < nb; g += (int64_t)gridDim.x * nw) {
        const int64_t s0 = g * 16;
        // inputs -> bf16 LDS (16 x kIn)
        for (int e = l; e < 16 * kIn; e += 64) {
            const int i = e / kIn, k = e % kIn;
            const float v = (s0 + i < nb) ? X[(s0 + i) * kIn + k] : 0.0f;
            act[0][i * kLda + k] = f2bf(v);
        }
        mlp_wave_sync();
        int cur = 0;
#pragma unroll
        for (int layer = 0; layer < 3; ++layer) {
            fv_f4 acc[8];
            if (layer == 0)
                wave_gemm_nt<kIn, 8>(act[cur], kLda, wl[0], ld1, acc);
            else
                wave_gemm_nt<kHid, 8>(act[cur], kLda, wl[layer], ldh, acc);
            const int nxt = cur ^ 1;
#pragma unroll
            for (int t = 0; t < 8; ++t) {
                const int n = t * 16 + col;
                const float bias = m.b[layer][n];
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int i = rq + v;
                    const float z = acc[t][v] + bias;
                    const unsigned short h = f2bf(silu(z));
                    act[nxt][i * kLda + n] = h;
                    if (zs && s0 + i < nb) {
                        zs[((int64_t)layer * nb + s0 + i) * kHid + n] = z;
                        hs[((int64_t)layer * nb + s0 + i) * kHid + n] = h;
                    }
                }
            }
            mlp_wave_sync();
            cur = nxt;
        }
        fv_f4 out[1];
        wave_gemm_nt<kHid, 1>(act[cur], kLda, wl[3], ldh, out);
        const float bias = m.b[3][col];
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int64_t s = s0 + rq + v;
            if (s < nb) Y[s * kOut + col] = out[0][v] + bias;
        }
        mlp_wave_sync();
    }
}

// Backward of the weighted MSE  L = (1/nb) sum_s sum_j w_j (Y_sj - T_sj)^2
// down to the layer errors, kept as bf16 (the MFMA inputs of the weight
// gradients):  dz[(L * nb + s) * 128 + n] for L = 0..2 (hidden layers), and
// dy[s * 16 + j] for the output layer.  STAGE as for the forward pass.
template <bool STAGE>
__global__ __launch_bounds__(STAGE ? 64 * kMlpWaves : 64) void k_mlp_backward(
    MlpW m, const float *__restrict__ Y, const float *__restrict__ T,
    const float *__restrict__ wout, int64_t nb, const float *__restrict__ zs,
    unsigned short *__restrict__ dz, unsigned short *__restrict__ dy)
{
    extern __shared__ __attribute__((aligned(16))) unsigned short smem[];
    const unsigned short *wt1 = m.wt[1], *wt2 = m.wt[2], *wt3 = m.wt[3];
    const int ldh = STAGE ? kLdwH : kHid, ldt = STAGE ? kLdwT : kOut;
    if (STAGE) {
        wt1 = smem + kBw1;
        wt2 = smem + kBw2;
        wt3 = smem + kBw3;
        stage_rows(m.wt[1], kHid, kHid, smem + kBw1, kLdwH);
        stage_rows(m.wt[2], kHid, kHid, smem + kBw2, kLdwH);
        stage_rows(m.wt[3], kHid, kOut, smem + kBw3, kLdwT);
        __syncthreads();
    }
    const int wv = threadIdx.x >> 6;
    unsigned short(*buf)[16 * kLda] = reinterpret_cast<unsigned short(*)[16 * kLda]>(
        smem + (STAGE ? kBwEnd : 0) + wv * kActPerWave);
    const int l = threadIdx.x & 63;
    const int col = l & 15, rq = (l >> 4) * 4;
    const float inv = 2.0f / (float)nb;
    const int nw = blockDim.x >> 6;
    for (int64_t g = (int64_t)blockIdx.x * nw + wv; g * 16 < nb; g += (int64_t)gridDim.x * nw) {
        const int64_t s0 = g * 16;
        for (int e = l; e < 256; e += 64) {
            const int i = e >> 4, j = e & 15;
            const int64_t s = s0 + i;
            float d = 0.0f;
            if (s < nb) d = inv * wout[j] * (Y[s * kOut + j] - T[s * kOut + j]);
            const unsigned short h = f2bf(d);
            buf[0][i * kLda + j] = h;
            if (s < nb) dy[s * kOut + j] = h;
        }
        mlp_wave_sync();
        int cur = 0;
        for (int layer = 2; layer >= 0; --layer) {
            fv_f4 acc[8];
            if (layer == 2)
                wave_gemm_nt<kOut, 8>(buf[cur], kLda, wt3, ldt, acc);
            else
                wave_gemm_nt<kHid, 8>(buf[cur], kLda, layer == 1 ? wt2 : wt1, ldh, acc);
            const int nxt = cur ^ 1;
#pragma unroll
            for (int t = 0; t < 8; ++t) {
                const int n = t * 16 + col;
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int i = rq + v;
                    const int64_t s = s0 + i;
                    float d = 0.0f;
                    if (s < nb) d = acc[t][v] * dsilu(zs[((int64_t)layer * nb + s) * kHid + n]);
                    const unsigned short h = f2bf(d);
                    buf[nxt][i * kLda + n] = h;
                    if (s < nb) dz[((int64_t)layer * nb + s) * kHid + n] = h;
                }
            }
            mlp_wave_sync();
            cur = nxt;
        }
    }
}

// Weight gradients dW[n][k] = sum_s E[s][n] H[s][k] and bias gradients
// db[n] = sum_s E[s][n] of all four layers, in two launches:
//   k_mlp_wgrad_part: one wavefront per (16 x 16 tile of any layer,
//     kWgChunk-sample chunk) accumulates the tile over its chunk on MFMA
//     (16-sample slices in order) and, for k-tile 0, the chunk's bias sums (4
//     sample lanes per column, combined in a fixed order); each chunk's
//     partials land in the parameter layout (part[chunk][kParams]);
//   k_mlp_wgrad_reduce: sums the chunks in order per parameter
//     (deterministic, no atomics).
constexpr int kWgChunk = 256;

struct WgLayer {
    const unsigned short *E;   // layer errors (bf16, nb x ne)
    const unsigned short *H;   // layer inputs (bf16, nb x nh), or null: bf16(X)
    const float *X;            // the network input (fp32, nb x nh) when H is null
    int ne, nh, off_w, off_b, tile0;
};
struct WgPlan {
    WgLayer L[4];
};

__global__ __launch_bounds__(64) void k_mlp_wgrad_part(WgPlan plan, int64_t nb,
                                                       float *__restrict__ part)
{
    const int l = threadIdx.x;
    int li = 0;
#pragma unroll
    for (int k = 1; k < 4; ++k)
        if ((int)blockIdx.x >= plan.L[k].tile0) li = k;
    const WgLayer &L = plan.L[li];
    const int ne = L.ne, nh = L.nh;
    const unsigned short *__restrict__ E = L.E;
    const unsigned short *__restrict__ H = L.H;
    const float *__restrict__ X = L.X;
    const int tile = (int)blockIdx.x - L.tile0;
    const int tiles_k = nh / 16;
    const int n0 = (tile / tiles_k) * 16, k0 = (tile % tiles_k) * 16;
    const int r = l & 15, sq = (l >> 4) * 4;
    const int64_t c0 = (int64_t)blockIdx.y * kWgChunk;
    const int64_t c1 = (c0 + kWgChunk < nb) ? c0 + kWgChunk : nb;
    float *pc = part + (int64_t)blockIdx.y * kParams;
    fv_f4 acc = {0.f, 0.f, 0.f, 0.f};
    for (int64_t s0 = c0; s0 < c1; s0 += 16) {
        fv_s4 a, bq;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int64_t s = s0 + sq + q;
            const bool ok = s < c1;
            a[q] = ok ? (short)E[s * ne + n0 + r] : (short)0;
            unsigned short hv = 0;
            if (ok) hv = H ? H[s * nh + k0 + r] : f2bf(X[s * nh + k0 + r]);
            bq[q] = (short)hv;
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, bq, acc, 0, 0, 0);
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) pc[L.off_w + (n0 + sq + v) * nh + k0 + r] = acc[v];
    if (k0 == 0) {
        // lane l: column n0 + (l & 15), samples c0 + (l >> 4) + 4 i
        float bs = 0.0f;
        for (int64_t s = c0 + (l >> 4); s < c1; s += 4) bs += bf2f(E[s * ne + n0 + r]);
        bs += __shfl_xor(bs, 16, 64);
        bs += __shfl_xor(bs, 32, 64);
        if (l < 16) pc[L.off_b + n0 + l] = bs;
    }
}

__global__ void k_mlp_wgrad_reduce(const float *__restrict__ part, int nchunk,
                                   float *__restrict__ g)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= kParams) return;
    float s = 0.0f;
    for (int c = 0; c < nchunk; ++c) s += part[(int64_t)c * kParams + i];
    g[i] = s;
}

// Adam on the FP32 master parameters; refreshes the bf16 weight copies.
__global__ void k_mlp_adam(float *__restrict__ p, const float *__restrict__ g,
                           float *__restrict__ mo, float *__restrict__ ve, int n, float lr,
                           float b1, float b2, float eps, float c1, float c2,
                           unsigned short *__restrict__ wbf, unsigned short *__restrict__ wtbf)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float m = b1 * mo[i] + (1.0f - b1) * gi;
    const float v = b2 * ve[i] + (1.0f - b2) * gi * gi;
    mo[i] = m;
    ve[i] = v;
    const float upd = lr * (m / c1) / (sqrtf(v / c2) + eps);
    const float np = p[i] - upd;
    p[i] = np;
    // bf16 copies of the weight matrices (biases stay fp32)
    int off = -1, rows = 0, cols = 0, base = 0;
    if (i >= kOffW1 && i < kOffB1) { off = i - kOffW1; rows = kHid; cols = kIn; base = 0; }
    else if (i >= kOffW2 && i < kOffB2) { off = i - kOffW2; rows = kHid; cols = kHid; base = 1; }
    else if (i >= kOffW3 && i < kOffB3) { off = i - kOffW3; rows = kHid; cols = kHid; base = 2; }
    else if (i >= kOffW4 && i < kOffB4) { off = i - kOffW4; rows = kOut; cols = kHid; base = 3; }
    if (off >= 0) {
        static constexpr int woff[4] = {0, kHid * kIn, kHid * kIn + kHid * kHid,
                                        kHid * kIn + 2 * kHid * kHid};
        const unsigned short h = f2bf(np);
        wbf[woff[base] + off] = h;
        const int r = off / cols, c = off % cols;
        wtbf[woff[base] + c * rows + r] = h;
    }
}

}  // namespace fv

using namespace fv;

struct fv_mlp {
    fv_ctx *ctx = nullptr;
    void *block = nullptr;
    float *params = nullptr, *grad = nullptr, *m1 = nullptr, *m2 = nullptr;
    unsigned short *wbf = nullptr, *wtbf = nullptr;
    int64_t step = 0;
    // training scratch (grown on demand)
    void *scratch = nullptr;
    int64_t scratch_nb = 0;
    // stream of the network's launches when set (a caller that orders them
    // against its own streams, mlp_set_stream), else the context's
    hipStream_t on = nullptr;
};

namespace fv {
static hipStream_t mlp_stream(const fv_mlp *m) { return m->on ? m->on : m->ctx->stream; }

// Sets the stream of the network's launches (nullptr: its context's);
// returns the previous setting.
hipStream_t mlp_set_stream(fv_mlp *m, hipStream_t s)
{
    const hipStream_t prev = m->on;
    m->on = s;
    return prev;
}
}  // namespace fv

static constexpr int kWeightsBf = kHid * kIn + 2 * kHid * kHid + kOut * kHid;

// Shape of a forward / backward launch.  Large batches (at least four
// 16-sample groups per CU): one workgroup per CU, kMlpWaves wavefronts
// sharing weights staged in LDS (~150 KB).  Small batches (the online
// training's few thousand samples): one wavefront per workgroup and group,
// the weights read from L2, since staging them per workgroup would cost more
// than the few groups each workgroup holds.
struct MlpShape {
    unsigned grid;
    int waves;
    bool stage;
    size_t lds_fwd, lds_bwd;
};
static MlpShape mlp_shape(fv_ctx *ctx, int64_t nb)
{
    const int64_t groups = ceil_div(nb, 16), cus = num_cus(ctx);
    if (groups >= 4 * cus) {
        const int64_t wg = ceil_div(groups, kMlpWaves);
        return MlpShape{(unsigned)(wg < cus ? wg : cus), kMlpWaves, true, fwd_lds(kMlpWaves),
                        bwd_lds(kMlpWaves)};
    }
    const size_t act = sizeof(unsigned short) * kActPerWave;
    return MlpShape{(unsigned)(groups < 8192 ? groups : 8192), 1, false, act, act};
}

static MlpW mlp_views(const fv_mlp *m)
{
    MlpW w;
    const int woff[4] = {0, kHid * kIn, kHid * kIn + kHid * kHid, kHid * kIn + 2 * kHid * kHid};
    const int boff[4] = {kOffB1, kOffB2, kOffB3, kOffB4};
    for (int k = 0; k < 4; ++k) {
        w.w[k] = m->wbf + woff[k];
        w.wt[k] = m->wtbf + woff[k];
        w.b[k] = m->params + boff[k];
    }
    return w;
}

static uint16_t host_f2bf(float f)
{
    uint32_t u;
    memcpy(&u, &f, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

static int mlp_scratch(fv_mlp *m, int64_t nb)
{
    if (m->scratch_nb >= nb) return FV_OK;
    fv_ctx *ctx = m->ctx;
    if (m->scratch) {
        FV_HIP_CHECK(ctx, hipStreamSynchronize(mlp_stream(m)));
        FV_HIP_CHECK(ctx, hipFree(m->scratch));
        m->scratch = nullptr;
    }
    // Y (16 fp32), zs (3x128 fp32), hs (3x128 bf16), dz (3x128 bf16), dy (16 bf16),
    // then the weight-gradient partials of every layer per sample chunk (parameter layout)
    const size_t nchunk = (size_t)ceil_div(nb, kWgChunk);
    const size_t bytes = (size_t)nb * (16 * 4 + 384 * 4 + 384 * 2 + 384 * 2 + 16 * 2) + 1024 +
                         nchunk * (size_t)kParams * sizeof(float) + 256;
    FV_HIP_CHECK(ctx, hipMalloc(&m->scratch, bytes));
    m->scratch_nb = nb;
    return FV_OK;
}

extern "C" {

int fv_mlp_create(fv_ctx *ctx, fv_mlp **out)
{
    if (!ctx || !out) return FV_E_ARG;
    *out = nullptr;
    fv_mlp *m = new (std::nothrow) fv_mlp();
    if (!m) return set_error(ctx, FV_E_HIP, "out of host memory");
    m->ctx = ctx;
    const size_t bytes = sizeof(float) * 4 * kParams + sizeof(unsigned short) * 2 * kWeightsBf + 256;
    hipError_t e = hipMalloc(&m->block, bytes);
    if (e != hipSuccess) {
        delete m;
        return set_error(ctx, FV_E_HIP, "hipMalloc for the MLP failed: %s", hipGetErrorString(e));
    }
    float *p = static_cast<float *>(m->block);
    m->params = p;
    m->grad = p + kParams;
    m->m1 = p + 2 * kParams;
    m->m2 = p + 3 * kParams;
    m->wbf = reinterpret_cast<unsigned short *>(p + 4 * kParams);
    m->wtbf = m->wbf + kWeightsBf;
    e = hipMemsetAsync(m->block, 0, bytes, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
        (void)hipFree(m->block);
        delete m;
        return set_error(ctx, FV_E_HIP, "MLP init failed: %s", hipGetErrorString(e));
    }
    *out = m;
    return FV_OK;
}

void fv_mlp_destroy(fv_mlp *m)
{
    if (!m) return;
    (void)hipStreamSynchronize(m->ctx->stream);
    if (m->scratch) (void)hipFree(m->scratch);
    (void)hipFree(m->block);
    delete m;
}

int fv_mlp_param_count(void) { return kParams; }

int fv_mlp_set_params(fv_mlp *m, const float *params)
{
    if (!m || !params) return FV_E_ARG;
    fv_ctx *ctx = m->ctx;
    std::vector<uint16_t> wb(kWeightsBf), wt(kWeightsBf);
    const int woff[4] = {0, kHid * kIn, kHid * kIn + kHid * kHid, kHid * kIn + 2 * kHid * kHid};
    const int poff[4] = {kOffW1, kOffW2, kOffW3, kOffW4};
    const int rows[4] = {kHid, kHid, kHid, kOut}, cols[4] = {kIn, kHid, kHid, kHid};
    for (int k = 0; k < 4; ++k)
        for (int r = 0; r < rows[k]; ++r)
            for (int c = 0; c < cols[k]; ++c) {
                const uint16_t h = host_f2bf(params[poff[k] + r * cols[k] + c]);
                wb[woff[k] + r * cols[k] + c] = h;
                wt[woff[k] + c * rows[k] + r] = h;
            }
    FV_HIP_CHECK(ctx, hipMemcpyAsync(m->params, params, sizeof(float) * kParams,
                                     hipMemcpyHostToDevice, ctx->stream));
    FV_HIP_CHECK(ctx, hipMemcpyAsync(m->wbf, wb.data(), 2 * kWeightsBf, hipMemcpyHostToDevice,
                                     ctx->stream));
    FV_HIP_CHECK(ctx, hipMemcpyAsync(m->wtbf, wt.data(), 2 * kWeightsBf, hipMemcpyHostToDevice,
                                     ctx->stream));
    FV_HIP_CHECK(ctx, hipMemsetAsync(m->m1, 0, sizeof(float) * 2 * kParams, ctx->stream));
    FV_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    m->step = 0;
    return FV_OK;
}

int fv_mlp_get_params(fv_mlp *m, float *params)
{
    if (!m || !params) return FV_E_ARG;
    return fv_memcpy_d2h(m->ctx, params, m->params, sizeof(float) * kParams);
}

int fv_mlp_forward_dev(fv_mlp *m, const float *X, int64_t nb, float *Y)
{
    if (!m) return FV_E_ARG;
    fv_ctx *ctx = m->ctx;
    FV_REQUIRE(ctx, nb >= 0, "negative batch");
    if (nb == 0) return FV_OK;
    FV_REQUIRE(ctx, X && Y, "null X/Y");
    const MlpShape sh = mlp_shape(ctx, nb);
    TimedLaunch tl(ctx, "mlp_forward", mlp_stream(m));
    if (sh.stage)
        hipLaunchKernelGGL(k_mlp_forward<true>, dim3(sh.grid), dim3(64 * sh.waves),
                           (unsigned)sh.lds_fwd, mlp_stream(m), mlp_views(m), X, nb, Y,
                           (float *)nullptr, (unsigned short *)nullptr);
    else
        hipLaunchKernelGGL(k_mlp_forward<false>, dim3(sh.grid), dim3(64 * sh.waves),
                           (unsigned)sh.lds_fwd, mlp_stream(m), mlp_views(m), X, nb, Y,
                           (float *)nullptr, (unsigned short *)nullptr);
    FV_LAUNCH_CHECK(ctx);
    tl.done();
    return FV_OK;
}

// Gradient of the weighted MSE into m->grad (device); loss not computed here.
static int mlp_gradients(fv_mlp *m, const float *X, const float *T, const float *wout,
                         int64_t nb)
{
    fv_ctx *ctx = m->ctx;
    int rc = mlp_scratch(m, nb);
    if (rc) return rc;
    char *s = static_cast<char *>(m->scratch);
    float *Y = reinterpret_cast<float *>(s);
    float *zs = Y + nb * 16;
    unsigned short *hs = reinterpret_cast<unsigned short *>(zs + nb * 384);
    unsigned short *dz = hs + nb * 384;
    unsigned short *dy = dz + nb * 384;
    const MlpShape sh = mlp_shape(ctx, nb);
    const MlpW w = mlp_views(m);
    TimedLaunch tl(ctx, "mlp_train", mlp_stream(m));
    if (sh.stage) {
        hipLaunchKernelGGL(k_mlp_forward<true>, dim3(sh.grid), dim3(64 * sh.waves),
                           (unsigned)sh.lds_fwd, mlp_stream(m), w, X, nb, Y, zs, hs);
        FV_LAUNCH_CHECK(ctx);
        hipLaunchKernelGGL(k_mlp_backward<true>, dim3(sh.grid), dim3(64 * sh.waves),
                           (unsigned)sh.lds_bwd, mlp_stream(m), w, Y, T, wout, nb, zs, dz, dy);
    } else {
        hipLaunchKernelGGL(k_mlp_forward<false>, dim3(sh.grid), dim3(64 * sh.waves),
                           (unsigned)sh.lds_fwd, mlp_stream(m), w, X, nb, Y, zs, hs);
        FV_LAUNCH_CHECK(ctx);
        hipLaunchKernelGGL(k_mlp_backward<false>, dim3(sh.grid), dim3(64 * sh.waves),
                           (unsigned)sh.lds_bwd, mlp_stream(m), w, Y, T, wout, nb, zs, dz, dy);
    }
    FV_LAUNCH_CHECK(ctx);
    const size_t base = (size_t)nb * (16 * 4 + 384 * 4 + 384 * 2 + 384 * 2 + 16 * 2) + 1024;
    float *part = reinterpret_cast<float *>(s + ((base + 255) & ~size_t(255)));
    const int nchunk = (int)ceil_div(nb, kWgChunk);
    // W1 / b1: E = dz[0], H = bf16(X);  W2 / b2: E = dz[1], H = h1;
    // W3 / b3: E = dz[2], H = h2;  W4 / b4: E = dy (16), H = h3
    WgPlan plan;
    plan.L[0] = WgLayer{dz, nullptr, X, kHid, kIn, kOffW1, kOffB1, 0};
    plan.L[1] = WgLayer{dz + nb * kHid, hs, nullptr, kHid, kHid, kOffW2, kOffB2, 0};
    plan.L[2] = WgLayer{dz + 2 * nb * kHid, hs + nb * kHid, nullptr, kHid, kHid, kOffW3, kOffB3, 0};
    plan.L[3] = WgLayer{dy, hs + 2 * nb * kHid, nullptr, kOut, kHid, kOffW4, kOffB4, 0};
    int tiles = 0;
    for (int k = 0; k < 4; ++k) {
        plan.L[k].tile0 = tiles;
        tiles += (plan.L[k].ne / 16) * (plan.L[k].nh / 16);
    }
    hipLaunchKernelGGL(k_mlp_wgrad_part, dim3((unsigned)tiles, (unsigned)nchunk), dim3(64), 0,
                       mlp_stream(m), plan, nb, part);
    FV_LAUNCH_CHECK(ctx);
    hipLaunchKernelGGL(k_mlp_wgrad_reduce, dim3((unsigned)ceil_div(kParams, 256)), dim3(256), 0,
                       mlp_stream(m), part, nchunk, m->grad);
    FV_LAUNCH_CHECK(ctx);
    tl.done();
    return FV_OK;
}

int fv_mlp_gradients_dev(fv_mlp *m, const float *X, const float *T, const float *wout,
                         int64_t nb, float *grad)
{
    if (!m) return FV_E_ARG;
    fv_ctx *ctx = m->ctx;
    FV_REQUIRE(ctx, nb > 0 && X && T && wout, "bad gradient arguments");
    int rc = mlp_gradients(m, X, T, wout, nb);
    if (rc) return rc;
    if (grad)
        FV_HIP_CHECK(ctx, hipMemcpyAsync(grad, m->grad, sizeof(float) * kParams,
                                         hipMemcpyDeviceToDevice, mlp_stream(m)));
    return FV_OK;
}

int fv_mlp_adam_dev(fv_mlp *m, const float *grad, float lr)
{
    if (!m) return FV_E_ARG;
    fv_ctx *ctx = m->ctx;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    m->step += 1;
    const float c1 = 1.0f - powf(b1, (float)m->step);
    const float c2 = 1.0f - powf(b2, (float)m->step);
    TimedLaunch tl(ctx, "mlp_adam", mlp_stream(m));
    hipLaunchKernelGGL(k_mlp_adam, dim3((unsigned)ceil_div(kParams, 256)), dim3(256), 0,
                       mlp_stream(m), m->params, grad ? grad : m->grad, m->m1, m->m2, kParams, lr,
                       b1, b2, eps, c1, c2, m->wbf, m->wtbf);
    FV_LAUNCH_CHECK(ctx);
    tl.done();
    return FV_OK;
}

int fv_mlp_train_step_dev(fv_mlp *m, const float *X, const float *T, const float *wout,
                          int64_t nb, float lr)
{
    if (!m) return FV_E_ARG;
    fv_ctx *ctx = m->ctx;
    FV_REQUIRE(ctx, nb > 0 && X && T && wout, "bad training arguments");
    int rc = mlp_gradients(m, X, T, wout, nb);
    if (rc) return rc;
    return fv_mlp_adam_dev(m, nullptr, lr);
}

// Host twins (tests): copy in, run, copy out, synchronise.
int fv_mlp_forward_host(fv_mlp *m, const float *X, int64_t nb, float *Y)
{
    if (!m) return FV_E_ARG;
    fv_ctx *ctx = m->ctx;
    FV_REQUIRE(ctx, nb >= 0 && (nb == 0 || (X && Y)), "bad forward arguments");
    if (nb == 0) return FV_OK;
    float *d = static_cast<float *>(workspace(ctx, S_TMP0, sizeof(float) * nb * (kIn + kOut)));
    if (!d) return FV_E_HIP;
    FV_HIP_CHECK(ctx, hipMemcpyAsync(d, X, sizeof(float) * nb * kIn, hipMemcpyHostToDevice,
                                     ctx->stream));
    int rc = fv_mlp_forward_dev(m, d, nb, d + nb * kIn);
    if (rc) return rc;
    return fv_memcpy_d2h(ctx, Y, d + nb * kIn, sizeof(float) * nb * kOut);
}

int fv_mlp_gradients_host(fv_mlp *m, const float *X, const float *T, const float *wout,
                          int64_t nb, float *grad)
{
    if (!m) return FV_E_ARG;
    fv_ctx *ctx = m->ctx;
    FV_REQUIRE(ctx, nb > 0 && X && T && wout && grad, "bad gradient arguments");
    float *d = static_cast<float *>(
        workspace(ctx, S_TMP1, sizeof(float) * (nb * (kIn + kOut) + kOut)));
    if (!d) return FV_E_HIP;
    FV_HIP_CHECK(ctx, hipMemcpyAsync(d, X, sizeof(float) * nb * kIn, hipMemcpyHostToDevice,
                                     ctx->stream));
    FV_HIP_CHECK(ctx, hipMemcpyAsync(d + nb * kIn, T, sizeof(float) * nb * kOut,
                                     hipMemcpyHostToDevice, ctx->stream));
    FV_HIP_CHECK(ctx, hipMemcpyAsync(d + nb * (kIn + kOut), wout, sizeof(float) * kOut,
                                     hipMemcpyHostToDevice, ctx->stream));
    int rc = mlp_gradients(m, d, d + nb * kIn, d + nb * (kIn + kOut), nb);
    if (rc) return rc;
    return fv_memcpy_d2h(ctx, grad, m->grad, sizeof(float) * kParams);
}

int fv_mlp_adam_host(fv_mlp *m, const float *grad, float lr)
{
    if (!m || !grad) return FV_E_ARG;
    fv_ctx *ctx = m->ctx;
    FV_HIP_CHECK(ctx, hipMemcpyAsync(m->grad, grad, sizeof(float) * kParams,
                                     hipMemcpyHostToDevice, ctx->stream));
    int rc = fv_mlp_adam_dev(m, nullptr, lr);
    if (rc) return rc;
    return fv_ctx_synchronize(ctx);
}

}  // extern "C"
