// fv_closure.hip -- learned far-field closure of the particle sums.
//
// The particles are binned into a uniform grid of cells (edge h).  For a
// target particle in cell t and a source cell s:
//   near (grid-centre distance <= kappa h): exact pair sums, LDS-tiled;
//   far: the source cell's first-order expansion about its centroid
//        (monopole A = Sum alpha, dipole M = Sum alpha (x - xbar)^T, FP64),
//        evaluated at every target particle;
//   far cells in the shell kappa h < distance <= 2 kappa h also carry a learned
//        residual (the quadrupole order and beyond): the closure MLP (BF16 on
//        MFMA) predicts it from the position relative to the cell and the
//        cell's normalised first and second moments, once per (target cell,
//        source cell) pair at the target cell's centre.
// The MLP is trained online on exact direct sums the solver generates (random
// shell pairs of the current state, one random target particle each), and the
// closure is checked against the exact direct sum every `check_every` steps:
// above `tol` (max-norm error of the particle-induced velocity relative to its
// max norm; a non-finite value counts as above) it is rejected and the direct
// sums are used until the next check passes.
//
// Beyond the cutoff the smoothing is inactive (g(rho) -> 1 for both kernels at
// rho = R / sigma >> 1), so one singular expansion serves the velocity,
// stretching and vorticity passes.
#include <hipcub/hipcub.hpp>

#include <cmath>
#include <cstring>
#include <new>
#include <vector>

#include "fv_vpm.h"

struct fv_mlp;
extern "C" int fv_mlp_forward_dev(fv_mlp *m, const float *X, int64_t nb, float *Y);
extern "C" int fv_mlp_train_step_dev(fv_mlp *m, const float *X, const float *T,
                                     const float *wout, int64_t nb, float lr);

namespace fv {

hipStream_t mlp_set_stream(fv_mlp *m, hipStream_t s);   // fv_mlp.hip

// The network's launches on stream s for the scope (the closure orders them
// against the field's stream itself).
struct MlpOn {
    fv_mlp *m;
    hipStream_t prev;
    MlpOn(fv_mlp *mm, hipStream_t s) : m(mm), prev(mm ? mlp_set_stream(mm, s) : nullptr) {}
    ~MlpOn()
    {
        if (m) mlp_set_stream(m, prev);
    }
    MlpOn(const MlpOn &) = delete;
    MlpOn &operator=(const MlpOn &) = delete;
};

// Per occupied cell: n, xbar(3), A(3), M(9, row-major M[r][k] = Sum a_r d_k),
// centre(3), S = Sum |alpha|, L2 = Sum |alpha| |d|^2 / S, Q(18: per component r
// the symmetric Sum a_r d_k d_l for kl = 00 01 02 11 12 22), pad.
constexpr int kMomStride = 40;
constexpr int kMomTile = 19;        // n .. centre: what the far-field pass stages
constexpr int kFeat = 32;           // MLP input width (31 features + pad)
constexpr int kOutW = 16;           // MLP output width (3 used)
constexpr double kShell = 2.0;      // residual shell: kappa h < R <= kShell kappa h
constexpr int kMaxSlices = 16;      // near (dx ranges) and far (source-cell ranges) slices
constexpr int kFarTile = 64;        // source cells staged per far-field tile
constexpr int kLocStride = 32;      // local expansion per cell: u0 (3), J0 (9), H (18), pad
constexpr int32_t kKeyEnd = 1 << 18;   // sort key of the slots past the particle count: above
                                       // every cell index (a grid has at most 262 144 cells)
constexpr int kKeyBits = 19;           // radix-sort bits covering kKeyEnd

// Binning parameters, computed on the device from the particles' bounding box
// and the counts that size the pair lists; the host reads them once per
// binning, after the last of those counts.
struct ClGrid {
    double ox, oy, oz, h;              // grid origin (bounding-box minimum), cell edge
    int64_t n;                         // particles binned (the device count)
    int64_t ncell, nocc, maxpop;       // cells, occupied cells, largest population
    int64_t nl, nb;                    // shell-pair and far-band list lengths
    int32_t nx, ny, nz, bad;           // bad: a non-finite position
};

// Cell count bound of a grid: cells <= min(8 max_cells, 262 144).
__host__ __device__ inline bool grid_fits(double cells, double max_cells)
{
    return cells <= max_cells * 8.0 && cells <= 262144.0;
}

// Cell edge h (from h0, grown by 1.25 until the grid fits) and grid size for
// the bounding box bb = (min xyz, max xyz), finite.
__host__ __device__ inline void grid_fit(const double *bb, double h, double max_cells, double *h_out,
                                         int *nx, int *ny, int *nz)
{
    for (;;) {
        const double gx = floor((bb[3] - bb[0]) / h) + 1, gy = floor((bb[4] - bb[1]) / h) + 1,
                     gz = floor((bb[5] - bb[2]) / h) + 1;
        if (grid_fits(gx * gy * gz, max_cells)) {
            *nx = (int)gx;
            *ny = (int)gy;
            *nz = (int)gz;
            *h_out = h;
            return;
        }
        h *= 1.25;
    }
}

// Squared cutoff (kappa h)^2 (1 + 1e-12) of the near field, the shell and the
// local expansions (the host and the device take the same expression).
__host__ __device__ inline double cut2(double kappa, double h)
{
    return (kappa * h) * (kappa * h) * (1.0 + 1e-12);
}

struct fv_closure_state {
    fv_mlp *mlp = nullptr;
    double h = 2.0, kappa = 4.0, tol = 1e-3, lr = 1e-3;
    double kappa_local = 0.0;          // > 0: cells beyond kappa_local h via local expansions
    int near_tpt[4][2][2] = {};        // targets per lane of k_near per (mode, kernel, precision)
    int check_every = 10, train_samples = 2048, train_iters = 2;
    int64_t max_cells = 1536;
    // status
    bool active = false;
    int64_t steps = 0, checks = 0, rejections = 0;
    double last_err = -1.0;
    double far_evals = 0.0;
    uint64_t seed = 0x9E3779B97F4A7C15ull;
    // geometry of the current build
    int64_t built_epoch = -1;
    int64_t bin_id = 0;                // bumped by every full binning
    int64_t binned_step = -1;          // closure step whose binning is current
    int64_t res_bin = -1, res_version = -1;
    int res_mlp = -1;
    int64_t mlp_version = 0;
    int64_t n = 0;                     // particles binned
    double ox = 0, oy = 0, oz = 0, hh = 2.0;
    int nx = 1, ny = 1, nz = 1;
    int64_t ncell = 0, nocc = 0, maxpop = 0, nlist = 0;
    int skx = -1, sky = -1, skz = -1;  // near-stencil half-widths in cells
    int nsl = 0;                       // near slices (contiguous dx ranges)
    std::vector<int> stencil;          // (dx, dy, dz_max) columns, dx-major
    std::vector<int> slice_off;        // first column of each slice, nsl + 1
    // device buffers
    void *block = nullptr;
    size_t block_bytes = 0;
    int32_t *key = nullptr, *key_s = nullptr, *idx = nullptr, *perm = nullptr;
    double *sx = nullptr, *sy = nullptr, *sz = nullptr, *sax = nullptr, *say = nullptr,
           *saz = nullptr, *ssg = nullptr;
    double *npart = nullptr;           // near partial sums, kMaxSlices x 3 per sorted particle
    double *fpart = nullptr;           // far partial sums, kMaxSlices x 3 per sorted particle
    void *cells = nullptr;
    size_t cells_bytes = 0;
    int32_t *occ = nullptr, *occ_of = nullptr;
    int32_t *pstart = nullptr;         // cell starts in sorted order, ncell + 1
    int32_t *cstart = nullptr, *cend = nullptr;     // pstart and pstart + 1: cell c's range
    void *far = nullptr;
    size_t far_bytes = 0;
    double *mom = nullptr;             // kMomStride per occupied cell
    double *res = nullptr;             // learned residual velocity, 3 per occupied cell
    double *loc = nullptr;             // local expansions, kLocStride per occupied cell
    void *lpart = nullptr;             // local-expansion partial sums per source chunk
    size_t lpart_bytes = 0;
    int32_t *lcnt = nullptr, *loff = nullptr;   // shell pairs per target cell, offsets
    void *band = nullptr;              // far band list (kappa h, kappa_local h] per target cell
    size_t band_bytes = 0;
    int32_t *bcnt = nullptr, *boff = nullptr, *bt = nullptr, *bs = nullptr;
    int64_t nband = 0;
    void *list = nullptr;
    size_t list_bytes = 0;
    int32_t *lt = nullptr, *ls = nullptr;       // shell pair list (target, source)
    double *lk = nullptr;              // residual scale per pair
    float *X = nullptr, *Y = nullptr;  // MLP inputs / outputs per pair
    int32_t *dstencil = nullptr;       // columns, then slice offsets
    void *tmp = nullptr;
    size_t tmp_bytes = 0;
    double *bbox = nullptr;            // per-block bounding-box partials, near-pair counts, grid
    unsigned long long *near_cnt = nullptr;   // [0] near pairs of the binning, [1] of every pass
    ClGrid *grid = nullptr;            // device grid record (in the bbox allocation)
    float *tX = nullptr, *tT = nullptr, *tw = nullptr;
    double *vdir = nullptr, *vclo = nullptr, *verr = nullptr;
    // online training on a side stream (default priority), overlapping the
    // step's first near- and far-field passes; the field's stream waits for
    // ev_train where the weights are next read or the binning is rewritten
    hipStream_t side = nullptr;
    hipEvent_t ev_bin = nullptr, ev_train = nullptr;
    bool train_pending = false;
};

}  // namespace fv

struct fv_closure : fv::fv_closure_state {};

namespace fv {


// ---------------------------------------------------------------- kernels
__global__ void k_bbox_partial(const int64_t *__restrict__ np_dev, const double *__restrict__ x,
                               const double *__restrict__ y, const double *__restrict__ z,
                               double *__restrict__ part)
{
    __shared__ double s[6][256];
    const int64_t n = *np_dev;
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
        const double v[3] = {x[i], y[i], z[i]};
        for (int c = 0; c < 3; ++c) {
            lo[c] = fmin(lo[c], v[c]);
            hi[c] = fmax(hi[c], v[c]);
        }
    }
    for (int c = 0; c < 3; ++c) {
        s[c][threadIdx.x] = lo[c];
        s[3 + c][threadIdx.x] = hi[c];
    }
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w)
            for (int c = 0; c < 3; ++c) {
                s[c][threadIdx.x] = fmin(s[c][threadIdx.x], s[c][threadIdx.x + w]);
                s[3 + c][threadIdx.x] = fmax(s[3 + c][threadIdx.x], s[3 + c][threadIdx.x + w]);
            }
        __syncthreads();
    }
    if (threadIdx.x < 6) part[blockIdx.x * 6 + threadIdx.x] = s[threadIdx.x][0];
}

// The bounding box (the blocks' partials, one per lane of one wavefront,
// nbb <= 64) and the grid of a binning.  With no particles, or a non-finite
// position (bad = 1), a one-cell grid that bins nothing.
__global__ __launch_bounds__(64) void k_grid(const double *__restrict__ part, int nbb,
                                             const int64_t *__restrict__ np_dev, double h0,
                                             double max_cells, ClGrid *__restrict__ g,
                                             unsigned long long *__restrict__ near_pairs)
{
    const int l = threadIdx.x;
    double bb[6];
    for (int c = 0; c < 6; ++c) {
        double v = l < nbb ? part[l * 6 + c] : (c < 3 ? INFINITY : -INFINITY);
        for (int off = 32; off > 0; off >>= 1) {
            const double o = __shfl_xor(v, off, 64);
            v = c < 3 ? fmin(v, o) : fmax(v, o);
        }
        bb[c] = v;
    }
    if (l != 0) return;
    *near_pairs = 0;   // the near-field pairs of this binning, counted after the read
    const int64_t n = *np_dev;
    bool fin = true;
    for (int c = 0; c < 6; ++c) fin = fin && isfinite(bb[c]);
    g->n = n;
    g->nocc = g->maxpop = g->nl = g->nb = 0;
    g->bad = (n > 0 && !fin) ? 1 : 0;
    if (n == 0 || !fin) {
        g->ox = g->oy = g->oz = 0.0;
        g->h = h0;
        g->nx = g->ny = g->nz = 1;
        g->ncell = 1;
        return;
    }
    double h;
    int nx, ny, nz;
    grid_fit(bb, h0, max_cells, &h, &nx, &ny, &nz);
    g->ox = bb[0];
    g->oy = bb[1];
    g->oz = bb[2];
    g->h = h;
    g->nx = nx;
    g->ny = ny;
    g->nz = nz;
    g->ncell = (int64_t)nx * ny * nz;
}

// Cell key of every particle slot i < nu (launch bound); slots past the
// particle count take kKeyEnd and sort after every particle.
__global__ void k_cell_keys(int64_t nu, const ClGrid *__restrict__ g, const double *__restrict__ x,
                            const double *__restrict__ y, const double *__restrict__ z,
                            int32_t *__restrict__ key, int32_t *__restrict__ idx)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nu) return;
    idx[i] = (int32_t)i;
    if (i >= g->n) {
        key[i] = kKeyEnd;
        return;
    }
    const double ih = 1.0 / g->h;
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    int ix = (int)floor((x[i] - g->ox) * ih), iy = (int)floor((y[i] - g->oy) * ih),
        iz = (int)floor((z[i] - g->oz) * ih);
    ix = ix < 0 ? 0 : (ix >= nx ? nx - 1 : ix);
    iy = iy < 0 ? 0 : (iy >= ny ? ny - 1 : iy);
    iz = iz < 0 ? 0 : (iz >= nz ? nz - 1 : iz);
    key[i] = (ix * ny + iy) * nz + iz;
}

// Start of every cell's particle range in the sorted keys, p[c] = the lower
// bound of c among the first n keys, for c = 0 .. ncb (ncb >= the grid's
// cell count, so p[ncb] = n): cell c holds the sorted particles
// [p[c], p[c + 1]).
__global__ void k_cell_bounds(int64_t ncb, const ClGrid *__restrict__ g,
                              const int32_t *__restrict__ key_s, int32_t *__restrict__ p)
{
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c > ncb) return;
    int64_t lo = 0, hi = g->n;
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if ((int64_t)key_s[mid] < c)
            lo = mid + 1;
        else
            hi = mid;
    }
    p[c] = (int32_t)lo;
}

__global__ void k_gather(int64_t nu, const ClGrid *__restrict__ g, const int32_t *__restrict__ perm,
                         const double *__restrict__ x,
                         const double *__restrict__ y, const double *__restrict__ z,
                         const double *__restrict__ ax, const double *__restrict__ ay,
                         const double *__restrict__ az, const double *__restrict__ sg,
                         double *__restrict__ sx, double *__restrict__ sy, double *__restrict__ sz,
                         double *__restrict__ sax, double *__restrict__ say,
                         double *__restrict__ saz, double *__restrict__ ssg)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nu || i >= g->n) return;
    const int32_t j = perm[i];
    sx[i] = x[j]; sy[i] = y[j]; sz[i] = z[j];
    sax[i] = ax[j]; say[i] = ay[j]; saz[i] = az[j];
    ssg[i] = sg[j];
}

// Ordered compaction of the occupied cells (one workgroup of 16 wavefronts,
// 1024 cells a round: ballots, the wavefronts' counts added in order); the
// count and the largest population into the grid record.
__global__ __launch_bounds__(1024) void k_occupied(ClGrid *__restrict__ g,
                                                   const int32_t *__restrict__ cstart,
                                                   const int32_t *__restrict__ cend,
                                                   int32_t *__restrict__ occ,
                                                   int32_t *__restrict__ occ_of)
{
    __shared__ int32_t wcnt[16];
    __shared__ int32_t maxpop;
    const int64_t ncell = g->ncell;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const unsigned long long below = lane ? (~0ull >> (64 - lane)) : 0ull;
    if (threadIdx.x == 0) maxpop = 0;
    int64_t base = 0;
    int32_t mp = 0;
    for (int64_t c0 = 0; c0 < ncell; c0 += 1024) {
        const int64_t c = c0 + threadIdx.x;
        const int pop = c < ncell ? cend[c] - cstart[c] : 0;
        mp = pop > mp ? pop : mp;
        const unsigned long long m = __ballot(pop > 0);
        if (lane == 0) wcnt[w] = __popcll(m);
        __syncthreads();
        int64_t at = base;
        int32_t tot = 0;
        for (int k = 0; k < 16; ++k) {
            at += k < w ? wcnt[k] : 0;
            tot += wcnt[k];
        }
        if (c < ncell) {
            if (pop > 0) {
                const int64_t o = at + __popcll(m & below);
                occ[o] = (int32_t)c;
                occ_of[c] = (int32_t)o;
            } else {
                occ_of[c] = -1;
            }
        }
        base += tot;
        __syncthreads();
    }
    atomicMax(&maxpop, mp);
    __syncthreads();
    if (threadIdx.x == 0) {
        g->nocc = base;
        g->maxpop = maxpop;
    }
}

// Moments of each occupied cell (one wavefront per cell, fixed order; the
// grid is launched for a bound on the occupied count).
__global__ __launch_bounds__(64) void k_moments(
    const ClGrid *__restrict__ g, const int32_t *__restrict__ occ, const int32_t *__restrict__ cstart,
    const int32_t *__restrict__ cend, const double *__restrict__ sx, const double *__restrict__ sy,
    const double *__restrict__ sz, const double *__restrict__ sax, const double *__restrict__ say,
    const double *__restrict__ saz, double *__restrict__ mom)
{
    const int64_t o = blockIdx.x;
    if (o >= g->nocc) return;
    const double ox = g->ox, oy = g->oy, oz = g->oz, h = g->h;
    const int ny = g->ny, nz = g->nz;
    const int c = occ[o];
    const int b = cstart[c], e = cend[c];
    const int l = threadIdx.x;
    double px = 0, py = 0, pz = 0;
    for (int i = b + l; i < e; i += 64) {
        px += sx[i];
        py += sy[i];
        pz += sz[i];
    }
    for (int off = 32; off > 0; off >>= 1) {
        px += __shfl_xor(px, off, 64);
        py += __shfl_xor(py, off, 64);
        pz += __shfl_xor(pz, off, 64);
    }
    const double inv = 1.0 / (double)(e - b);
    const double xb = px * inv, yb = py * inv, zb = pz * inv;
    // acc: A(3) M(9) Q(18) S L2
    double acc[32];
    for (int k = 0; k < 32; ++k) acc[k] = 0.0;
    for (int i = b + l; i < e; i += 64) {
        const double a[3] = {sax[i], say[i], saz[i]};
        const double d[3] = {sx[i] - xb, sy[i] - yb, sz[i] - zb};
        const double am = sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
        const double dd[6] = {d[0] * d[0], d[0] * d[1], d[0] * d[2],
                              d[1] * d[1], d[1] * d[2], d[2] * d[2]};
        for (int r = 0; r < 3; ++r) {
            acc[r] += a[r];
            for (int k = 0; k < 3; ++k) acc[3 + 3 * r + k] += a[r] * d[k];
            for (int q = 0; q < 6; ++q) acc[12 + 6 * r + q] += a[r] * dd[q];
        }
        acc[30] += am;
        acc[31] += am * (dd[0] + dd[3] + dd[5]);
    }
    for (int k = 0; k < 32; ++k)
        for (int off = 32; off > 0; off >>= 1) acc[k] += __shfl_xor(acc[k], off, 64);
    if (l == 0) {
        double *m = mom + o * kMomStride;
        m[0] = (double)(e - b);
        m[1] = xb; m[2] = yb; m[3] = zb;
        for (int r = 0; r < 12; ++r) m[4 + r] = acc[r];
        const int ix = c / (ny * nz), iy = (c / nz) % ny, iz = c % nz;
        m[16] = ox + (ix + 0.5) * h;
        m[17] = oy + (iy + 0.5) * h;
        m[18] = oz + (iz + 0.5) * h;
        m[19] = acc[30];
        m[20] = acc[30] > 0.0 ? acc[31] / acc[30] : 0.0;
        for (int q = 0; q < 18; ++q) m[21 + q] = acc[12 + q];
        m[39] = 0.0;
    }
}

// First-order (monopole + dipole) far field of a cell at point x, FP64:
//   u = -(1/4pi) d x A / R^3 + (1/4pi) [ e / R^3 - 3 d x (M d) / R^5 ],
//   e_i = eps_ikl M_lk,  d = x - xbar;  J = du/dx (when J is non-null).
// m holds n, xbar(3), A(3), M(9) at offsets 0..15.
__device__ void far_expansion(const double *m, double x0, double x1, double x2, double *u,
                              double *J)
{
    const double d[3] = {x0 - m[1], x1 - m[2], x2 - m[3]};
    const double *A = m + 4, *M = m + 7;
    const double R2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    const double iR = rsqrt_refined(R2);
    const double iR2 = iR * iR;
    const double iR3 = iR * iR2, iR5 = iR3 * iR2, iR7 = iR5 * iR2;
    const double c = kInv4Pi;
    const double dA[3] = {d[1] * A[2] - d[2] * A[1], d[2] * A[0] - d[0] * A[2],
                          d[0] * A[1] - d[1] * A[0]};
    const double e[3] = {M[7] - M[5], M[2] - M[6], M[3] - M[1]};
    const double Md[3] = {M[0] * d[0] + M[1] * d[1] + M[2] * d[2],
                          M[3] * d[0] + M[4] * d[1] + M[5] * d[2],
                          M[6] * d[0] + M[7] * d[1] + M[8] * d[2]};
    const double dMd[3] = {d[1] * Md[2] - d[2] * Md[1], d[2] * Md[0] - d[0] * Md[2],
                           d[0] * Md[1] - d[1] * Md[0]};
    for (int i = 0; i < 3; ++i) u[i] += c * (-dA[i] * iR3 + e[i] * iR3 - 3.0 * dMd[i] * iR5);
    if (!J) return;
    for (int i = 0; i < 3; ++i)
        for (int k = 0; k < 3; ++k) {
            const int i1 = (i + 1) % 3, i2 = (i + 2) % 3;
            double epsA = 0.0;   // eps_{i k l} A_l
            if (k == i1) epsA = A[i2];
            else if (k == i2) epsA = -A[i1];
            const double t0 = -(epsA * iR3 - 3.0 * dA[i] * d[k] * iR5);
            const double t1 = -3.0 * e[i] * d[k] * iR5;
            // d/dd_k (d x Md)_i = eps_ikn Md_n + eps_imn d_m M_nk
            double epsMd = 0.0;
            if (k == i1) epsMd = Md[i2];
            else if (k == i2) epsMd = -Md[i1];
            const double colk[3] = {M[k], M[3 + k], M[6 + k]};
            const double dxc = (i == 0) ? d[1] * colk[2] - d[2] * colk[1]
                               : (i == 1) ? d[2] * colk[0] - d[0] * colk[2]
                                          : d[0] * colk[1] - d[1] * colk[0];
            const double t2 = -3.0 * ((epsMd + dxc) * iR5 - 5.0 * dMd[i] * d[k] * iR7);
            J[3 * i + k] += c * (t0 + t1 + t2);
        }
}

// Local (target-cell) expansion of the far field of every occupied cell whose
// centre lies beyond kappa_local h from the target cell's centre.  With phi =
// 1/R and d = c_t - xbar_s, a cell's first-order field is
//   u_i = -c eps_ijk (A_j dk phi - M_jl dl dk phi),
// so u0, J0_im = dm u_i and H_imn = dn dm u_i at c_t follow from the
// derivatives of phi up to the fourth (contracted with M through v = M d, the
// eps sums written as cross products); the particles of the target cell then
// take u0 + J0 (x - c_t) + H (x - c_t)(x - c_t) / 2 (and J0 + H (x - c_t) for
// the gradient).  k_local_part: one target cell per lane, the source cells
// of one chunk per workgroup, every fourth one per wavefront (the same source
// for every lane of a wavefront: scalar loads), the four wavefronts' sums
// added in order through LDS: 30 partial coefficients per (chunk, target);
// k_local_sum adds the chunks in order.
__global__ __launch_bounds__(256) void k_local_part(int64_t nocc, int64_t chunk,
                                                    const double *__restrict__ mom, double rl2,
                                                    double *__restrict__ lpart)
{
    __shared__ double red[3][30][64];
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int64_t t = (int64_t)blockIdx.x * 64 + lane;
    const bool act = t < nocc;
    const double *mt = mom + (act ? t : nocc - 1) * kMomStride;
    const double ct[3] = {mt[16], mt[17], mt[18]};
    double acc[30];
    for (int q = 0; q < 30; ++q) acc[q] = 0.0;
    const double c = kInv4Pi;
    const int64_t s0 = (int64_t)blockIdx.y * chunk;
    const int64_t s1 = (s0 + chunk < nocc) ? s0 + chunk : nocc;
    for (int64_t s = s0 + w; s < s1; s += 4) {
        const double *ms = mom + s * kMomStride;
        const double gx = ct[0] - ms[16], gy = ct[1] - ms[17], gz = ct[2] - ms[18];
        if (gx * gx + gy * gy + gz * gz <= rl2) continue;
        const double d[3] = {ct[0] - ms[1], ct[1] - ms[2], ct[2] - ms[3]};
        const double *A = ms + 4, *M = ms + 7;   // M[3 j + l] = Sum a_j s_l
        const double R2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
        const double iR = rsqrt_refined(R2);
        const double iR2 = iR * iR;
        const double iR3 = iR * iR2, iR5 = iR3 * iR2, iR7 = iR5 * iR2, iR9 = iR7 * iR2;
        // v = M d; with (j1, k1) = (i+1, i+2) every eps_ijk sum becomes a
        // cross product or an antisymmetric pair
        const double v[3] = {M[0] * d[0] + M[1] * d[1] + M[2] * d[2],
                             M[3] * d[0] + M[4] * d[1] + M[5] * d[2],
                             M[6] * d[0] + M[7] * d[1] + M[8] * d[2]};
        const double Ad[3] = {A[1] * d[2] - A[2] * d[1], A[2] * d[0] - A[0] * d[2],
                              A[0] * d[1] - A[1] * d[0]};
        const double vd[3] = {v[1] * d[2] - v[2] * d[1], v[2] * d[0] - v[0] * d[2],
                              v[0] * d[1] - v[1] * d[0]};
        const double em[3] = {M[5] - M[7], M[6] - M[2], M[1] - M[3]};   // M_j1k1 - M_k1j1
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            const int j1 = (i + 1) % 3, k1 = (i + 2) % 3;
            acc[i] += -c * (-iR3 * Ad[i] - (3.0 * iR5 * vd[i] - iR3 * em[i]));
            double Y[3], Am[3], Vm[3];
#pragma unroll
            for (int m = 0; m < 3; ++m) {
                Y[m] = M[3 * j1 + m] * d[k1] - M[3 * k1 + m] * d[j1];
                Am[m] = (k1 == m ? A[j1] : 0.0) - (j1 == m ? A[k1] : 0.0);
                Vm[m] = (k1 == m ? v[j1] : 0.0) - (j1 == m ? v[k1] : 0.0);
            }
#pragma unroll
            for (int m = 0; m < 3; ++m) {
                const double t2 = 3.0 * iR5 * Ad[i] * d[m] - iR3 * Am[m];
                const double s3 = -15.0 * iR7 * vd[i] * d[m] + 3.0 * iR5 * (em[i] * d[m] + Y[m] + Vm[m]);
                acc[3 + 3 * i + m] += -c * (t2 - s3);
            }
            int o = 0;
#pragma unroll
            for (int m = 0; m < 3; ++m)
#pragma unroll
                for (int n = m; n < 3; ++n, ++o) {
                    const double dmn = m == n ? 1.0 : 0.0;
                    const double t3 = -15.0 * iR7 * Ad[i] * d[m] * d[n] +
                                      3.0 * iR5 * (Am[m] * d[n] + Am[n] * d[m] + dmn * Ad[i]);
                    const double z = (k1 == n ? M[3 * j1 + m] : 0.0) - (j1 == n ? M[3 * k1 + m] : 0.0) +
                                     (k1 == m ? M[3 * j1 + n] : 0.0) - (j1 == m ? M[3 * k1 + n] : 0.0);
                    const double s4 = 105.0 * iR9 * vd[i] * d[m] * d[n] -
                                      15.0 * iR7 * (em[i] * d[m] * d[n] + Y[m] * d[n] + Y[n] * d[m] +
                                                    Vm[m] * d[n] + Vm[n] * d[m] + dmn * vd[i]) +
                                      3.0 * iR5 * (em[i] * dmn + z);
                    acc[12 + 6 * i + o] += -c * (t3 - s4);
                }
        }
    }
    if (w > 0)
        for (int q = 0; q < 30; ++q) red[w - 1][q][lane] = acc[q];
    __syncthreads();
    if (w == 0 && act) {
        double *o = lpart + ((int64_t)blockIdx.y * nocc + t) * 30;
        for (int q = 0; q < 30; ++q) o[q] = ((acc[q] + red[0][q][lane]) + red[1][q][lane]) + red[2][q][lane];
    }
}

// loc[t][q] = sum over chunks (in chunk order) of lpart[chunk][t][q]; pads 0.
__global__ void k_local_sum(int64_t nocc, int nchunk, const double *__restrict__ lpart,
                            double *__restrict__ loc)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nocc * kLocStride) return;
    const int64_t t = e / kLocStride;
    const int q = (int)(e % kLocStride);
    double v = 0.0;
    if (q < 30)
        for (int k = 0; k < nchunk; ++k) v += lpart[((int64_t)k * nocc + t) * 30 + q];
    loc[e] = v;
}

// Adds the local expansion of the particle's cell (offset dd from the centre)
// for mode 0 (velocity), 2 (stretching J alpha) or 3 (vorticity).
template <int MODE>
__device__ __forceinline__ void local_term(const double *L, const double *dd, const double *al,
                                           double *r)
{
    double J[9];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const double *H = L + 12 + 6 * i;   // mn = 00 01 02 11 12 22
        const double Hd[3] = {H[0] * dd[0] + H[1] * dd[1] + H[2] * dd[2],
                              H[1] * dd[0] + H[3] * dd[1] + H[4] * dd[2],
                              H[2] * dd[0] + H[4] * dd[1] + H[5] * dd[2]};
        if (MODE == 0) {
            double v = L[i];
            for (int m = 0; m < 3; ++m) v += (L[3 + 3 * i + m] + 0.5 * Hd[m]) * dd[m];
            r[i] += v;
        } else {
            for (int m = 0; m < 3; ++m) J[3 * i + m] = L[3 + 3 * i + m] + Hd[m];
        }
    }
    if (MODE == 2)
        for (int i = 0; i < 3; ++i) r[i] += J[3 * i] * al[0] + J[3 * i + 1] * al[1] + J[3 * i + 2] * al[2];
    if (MODE == 3) {
        r[0] += J[7] - J[5];
        r[1] += J[2] - J[6];
        r[2] += J[3] - J[1];
    }
}

// The far term of one cell at point x for the closure passes, in closed form:
// with d = x - xbar, w = 3 (d x A - e) / R^5 + 15 d x (M d) / R^7,
//   J_ik = c ( -eps_ikl A_l / R^3 - 3 (eps_ikn (Md)_n + (d x M_.k)_i) / R^5 + w_i d_k ),
//   MODE 0: u (as far_expansion);  MODE 2: S = J alpha
//     = c ( -(alpha x A) / R^3 - 3 (alpha x Md + d x (M alpha)) / R^5 + w (d . alpha) );
//   MODE 3: omega_i = eps_ijk J_kj.
template <int MODE>
__device__ __forceinline__ void far_term(const double *m, double x0, double x1, double x2,
                                         const double *al, double *acc)
{
    const double d[3] = {x0 - m[1], x1 - m[2], x2 - m[3]};
    const double *A = m + 4, *M = m + 7;
    const double R2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    const double iR = rsqrt_refined(R2);
    const double iR2 = iR * iR;
    const double iR3 = iR * iR2, iR5 = iR3 * iR2;
    const double c = kInv4Pi;
    const double dA[3] = {d[1] * A[2] - d[2] * A[1], d[2] * A[0] - d[0] * A[2],
                          d[0] * A[1] - d[1] * A[0]};
    const double e[3] = {M[7] - M[5], M[2] - M[6], M[3] - M[1]};
    const double Md[3] = {M[0] * d[0] + M[1] * d[1] + M[2] * d[2],
                          M[3] * d[0] + M[4] * d[1] + M[5] * d[2],
                          M[6] * d[0] + M[7] * d[1] + M[8] * d[2]};
    const double dMd[3] = {d[1] * Md[2] - d[2] * Md[1], d[2] * Md[0] - d[0] * Md[2],
                           d[0] * Md[1] - d[1] * Md[0]};
    if (MODE == 0) {
        for (int i = 0; i < 3; ++i) acc[i] += c * ((e[i] - dA[i]) * iR3 - 3.0 * dMd[i] * iR5);
        return;
    }
    const double iR7 = iR5 * iR2;
    double w[3];
    for (int i = 0; i < 3; ++i) w[i] = 3.0 * (dA[i] - e[i]) * iR5 + 15.0 * dMd[i] * iR7;
    if (MODE == 2) {
        const double aA[3] = {al[1] * A[2] - al[2] * A[1], al[2] * A[0] - al[0] * A[2],
                              al[0] * A[1] - al[1] * A[0]};
        const double aMd[3] = {al[1] * Md[2] - al[2] * Md[1], al[2] * Md[0] - al[0] * Md[2],
                               al[0] * Md[1] - al[1] * Md[0]};
        const double Ma[3] = {M[0] * al[0] + M[1] * al[1] + M[2] * al[2],
                              M[3] * al[0] + M[4] * al[1] + M[5] * al[2],
                              M[6] * al[0] + M[7] * al[1] + M[8] * al[2]};
        const double dMa[3] = {d[1] * Ma[2] - d[2] * Ma[1], d[2] * Ma[0] - d[0] * Ma[2],
                               d[0] * Ma[1] - d[1] * Ma[0]};
        const double da = d[0] * al[0] + d[1] * al[1] + d[2] * al[2];
        for (int i = 0; i < 3; ++i)
            acc[i] += c * (-aA[i] * iR3 - 3.0 * (aMd[i] + dMa[i]) * iR5 + w[i] * da);
        return;
    }
    // MODE 3: J from the closed form, then its antisymmetric part
    double J[9];
    // -eps_ikl A_l / R^3 - 3 eps_ikn Md_n / R^5 = -eps_ikl (A_l / R^3 + 3 Md_l / R^5)
    const double v[3] = {A[0] * iR3 + 3.0 * Md[0] * iR5, A[1] * iR3 + 3.0 * Md[1] * iR5,
                         A[2] * iR3 + 3.0 * Md[2] * iR5};
    J[0] = 0.0;   J[1] = -v[2]; J[2] = v[1];
    J[3] = v[2];  J[4] = 0.0;   J[5] = -v[0];
    J[6] = -v[1]; J[7] = v[0];  J[8] = 0.0;
    for (int k = 0; k < 3; ++k) {
        const double col[3] = {M[k], M[3 + k], M[6 + k]};
        const double dc[3] = {d[1] * col[2] - d[2] * col[1], d[2] * col[0] - d[0] * col[2],
                              d[0] * col[1] - d[1] * col[0]};
        for (int i = 0; i < 3; ++i) J[3 * i + k] += -3.0 * dc[i] * iR5 + w[i] * d[k];
    }
    acc[0] += c * (J[7] - J[5]);
    acc[1] += c * (J[2] - J[6]);
    acc[2] += c * (J[3] - J[1]);
}

// MLP features of a (point, cell) pair:
//   [d/R (3), L/R, M/(S L) (9), Q/(S L^2) (18), 0]
// with S = Sum |alpha| and L^2 = max(Sum |alpha| |d|^2 / S, 1e-6 h^2); every
// moment feature lies in [-1, 1] (Cauchy-Schwarz).  Returns the residual scale
// k = S L^2 / (4 pi R^4): the network's output y (3) stands for the velocity
// residual k y.  A cell with S = 0 gets zero features and k = 0.
__device__ double far_features(const double *m, double x0, double x1, double x2, double h,
                               float *f)
{
    const double d[3] = {x0 - m[1], x1 - m[2], x2 - m[3]};
    const double R2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    const double R = sqrt(R2);
    const double S = m[19];
    if (!(S > 0.0)) {
        for (int k = 0; k < kFeat; ++k) f[k] = 0.0f;
        return 0.0;
    }
    const double L2 = fmax(m[20], 1e-6 * h * h);
    const double L = sqrt(L2);
    for (int c = 0; c < 3; ++c) f[c] = (float)(d[c] / R);
    f[3] = (float)(L / R);
    const double sm = 1.0 / (S * L), sq = 1.0 / (S * L2);
    for (int r = 0; r < 9; ++r) f[4 + r] = (float)(m[7 + r] * sm);
    for (int q = 0; q < 18; ++q) f[13 + q] = (float)(m[21 + q] * sq);
    f[31] = 0.0f;
    return S * L2 * kInv4Pi / (R2 * R2);
}

__device__ __forceinline__ double centre_dist2(const double *mt, const double *ms)
{
    const double dx = mt[16] - ms[16], dy = mt[17] - ms[17], dz = mt[18] - ms[18];
    return dx * dx + dy * dy + dz * dz;
}

// Pairs per target cell of the shell kappa h < D <= kShell kappa h and, when
// bcnt is non-null, of the band kappa h < D <= kappa_local h, D the distance
// of the cell centres: one wavefront per target cell, the grid launched for a
// bound nb - 1 on the occupied count; the counts are 0 for nocc <= t < nb (the
// trailing entries of the exclusive scans that follow).
__global__ __launch_bounds__(64) void k_shell_count(const ClGrid *__restrict__ g,
                                                    const double *__restrict__ mom, double kappa,
                                                    double kappa_local, int32_t *__restrict__ lcnt,
                                                    int32_t *__restrict__ bcnt)
{
    const int64_t t = blockIdx.x;
    const int64_t nocc = g->nocc;
    if (t >= nocc) {
        if (threadIdx.x == 0) {
            lcnt[t] = 0;
            if (bcnt) bcnt[t] = 0;
        }
        return;
    }
    const double rc2 = cut2(kappa, g->h);
    const double rm2 = rc2 * kShell * kShell;
    const double rl2 = bcnt ? cut2(kappa_local, g->h) : 0.0;
    const double *mt = mom + t * kMomStride;
    int c = 0, cb = 0;
    for (int64_t s = threadIdx.x; s < nocc; s += 64) {
        const double d2 = centre_dist2(mt, mom + s * kMomStride);
        c += (d2 > rc2 && d2 <= rm2) ? 1 : 0;
        cb += (d2 > rc2 && d2 <= rl2) ? 1 : 0;
    }
    for (int off = 32; off > 0; off >>= 1) {
        c += __shfl_xor(c, off, 64);
        cb += __shfl_xor(cb, off, 64);
    }
    if (threadIdx.x == 0) {
        lcnt[t] = c;
        if (bcnt) bcnt[t] = cb;
    }
}

// Exclusive scans of the shell counts and (bc non-null) the band counts by
// one workgroup of 16 wavefronts (4 consecutive entries a lane, 4096 a
// round) over the entries 0 .. nocc (the offsets the lists are read at; the
// counts past nocc are 0); the totals, the list lengths, into the grid record.
__global__ __launch_bounds__(1024) void k_scan_lists(ClGrid *__restrict__ g,
                                                     const int32_t *__restrict__ lc,
                                                     int32_t *__restrict__ lo,
                                                     const int32_t *__restrict__ bc,
                                                     int32_t *__restrict__ bo)
{
    __shared__ int32_t wsum[16];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int64_t nb = g->nocc + 1;
    for (int a = 0; a < (bc ? 2 : 1); ++a) {
        const int32_t *cnt = a ? bc : lc;
        int32_t *off = a ? bo : lo;
        int64_t carry = 0;
        for (int64_t base = 0; base < nb; base += 4096) {
            const int64_t i0 = base + 4 * (int64_t)threadIdx.x;
            int32_t v[4], run = 0;
            for (int q = 0; q < 4; ++q) {
                v[q] = i0 + q < nb ? cnt[i0 + q] : 0;
                run += v[q];
            }
            int32_t x = run;   // inclusive scan of the lanes' totals over the wavefront
            for (int d = 1; d < 64; d <<= 1) {
                const int32_t y = __shfl_up(x, d, 64);
                if (lane >= d) x += y;
            }
            if (lane == 63) wsum[w] = x;
            __syncthreads();
            int64_t pre = carry, tot = 0;
            for (int k = 0; k < 16; ++k) {
                pre += k < w ? wsum[k] : 0;
                tot += wsum[k];
            }
            int64_t o = pre + x - run;
            for (int q = 0; q < 4; ++q) {
                if (i0 + q < nb) off[i0 + q] = (int32_t)o;
                o += v[q];
            }
            carry += tot;
            __syncthreads();
        }
        if (threadIdx.x == 0) {
            if (a)
                g->nb = carry;
            else
                g->nl = carry;
        }
    }
}

// Ordered fill of the shell pair list (lt non-null: rc2 < D^2 <= rm2) and of
// the far band list (bt non-null: rc2 < D^2 <= rl2), source order within each
// target cell, at the offsets of the scans.
__global__ __launch_bounds__(64) void k_shell_fill(int64_t nocc, const double *__restrict__ mom,
                                                   double rc2, double rm2, double rl2,
                                                   const int32_t *__restrict__ loff,
                                                   int32_t *__restrict__ lt,
                                                   int32_t *__restrict__ ls,
                                                   const int32_t *__restrict__ boff,
                                                   int32_t *__restrict__ bt,
                                                   int32_t *__restrict__ bs)
{
    const int64_t t = blockIdx.x;
    if (t >= nocc) return;
    const double *mt = mom + t * kMomStride;
    const int lane = threadIdx.x;
    const unsigned long long below = lane ? (~0ull >> (64 - lane)) : 0ull;
    int64_t lb = lt ? loff[t] : 0, bb = bt ? boff[t] : 0;
    for (int64_t s0 = 0; s0 < nocc; s0 += 64) {
        const int64_t s = s0 + lane;
        bool inl = false, inb = false;
        if (s < nocc) {
            const double d2 = centre_dist2(mt, mom + s * kMomStride);
            inl = lt && d2 > rc2 && d2 <= rm2;
            inb = bt && d2 > rc2 && d2 <= rl2;
        }
        const unsigned long long ml = __ballot(inl), mb = __ballot(inb);
        if (inl) {
            const int64_t p = lb + __popcll(ml & below);
            lt[p] = (int32_t)t;
            ls[p] = (int32_t)s;
        }
        if (inb) {
            const int64_t p = bb + __popcll(mb & below);
            bt[p] = (int32_t)t;
            bs[p] = (int32_t)s;
        }
        lb += __popcll(ml);
        bb += __popcll(mb);
    }
}

// MLP inputs and residual scales of the shell pairs, at the target cell's centre.
__global__ void k_shell_inputs(int64_t nlist, const int32_t *__restrict__ lt,
                               const int32_t *__restrict__ ls, const double *__restrict__ mom,
                               double h, float *__restrict__ X, double *__restrict__ lk)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nlist) return;
    const double *mt = mom + (int64_t)lt[e] * kMomStride, *ms = mom + (int64_t)ls[e] * kMomStride;
    float f[kFeat];
    lk[e] = far_features(ms, mt[16], mt[17], mt[18], h, f);
    for (int k = 0; k < kFeat; ++k) X[e * kFeat + k] = f[k];
}

// Learned residual velocity per target cell: Sum over its shell pairs of k y
// (one wavefront per target cell, fixed order).
__global__ __launch_bounds__(64) void k_shell_reduce(int64_t nocc, const int32_t *__restrict__ off,
                                                     const double *__restrict__ lk,
                                                     const float *__restrict__ Y,
                                                     double *__restrict__ res)
{
    const int64_t t = blockIdx.x;
    if (t >= nocc) return;
    double a[3] = {0.0, 0.0, 0.0};
    for (int64_t e = off[t] + threadIdx.x; e < off[t + 1]; e += 64) {
        const double k = lk[e];
        for (int c = 0; c < 3; ++c) a[c] += k * (double)Y[e * kOutW + c];
    }
    for (int c = 0; c < 3; ++c)
        for (int o = 32; o > 0; o >>= 1) a[c] += __shfl_xor(a[c], o, 64);
    if (threadIdx.x == 0)
        for (int c = 0; c < 3; ++c) res[3 * t + c] = a[c];
}


// Near-field pair sums in sorted order.  The particles are sorted by cell key
// (ix * ny + iy) * nz + iz, so the cells of one stencil column (dx, dy, dz in
// [-dz_max, dz_max]) hold one contiguous particle range [pstart[k0],
// pstart[k1 + 1]).  One wavefront per (occupied target cell, 64 TPT-target
// chunk, dx slice), TPT targets per lane (1, 2 or 4; timed at first use when
// the context autotunes, else 2), sources streamed 64 at a time through LDS;
// each slice writes its own partial sums (summed in order later).
// At one target per lane a chunk of p <= 32 targets (a cell's last, partly
// filled chunk) runs as g = 64 / p lane groups of p: group q takes the tile's
// sources q, q + g, q + 2g, ..., and the groups' sums are added in group order
// before finish(), so the wavefront is not left mostly idle.
template <int KERNEL, int PREC, int MODE, int TPT>
__global__ __launch_bounds__(64) void k_near(
    int64_t n, int64_t nocc, const int32_t *__restrict__ occ, const int32_t *__restrict__ cstart,
    const int32_t *__restrict__ cend, const int32_t *__restrict__ pstart,
    const int32_t *__restrict__ cols, const int32_t *__restrict__ slice_off, int nx, int ny,
    int nz, const double *__restrict__ sx, const double *__restrict__ sy,
    const double *__restrict__ sz, const double *__restrict__ sax, const double *__restrict__ say,
    const double *__restrict__ saz, const double *__restrict__ ssg, double *__restrict__ npart)
{
    using Op = PartOp<KERNEL, PREC, MODE>;
    __shared__ PartRec tile[64];
    const int64_t o = blockIdx.x;
    if (o >= nocc) return;
    const int c = occ[o];
    const int tb = cstart[c], te = cend[c];
    const int t0 = tb + (int)blockIdx.y * (64 * TPT);
    if (t0 >= te) return;
    // lane groups (TPT == 1 only; wave-uniform p and g)
    const int p = (TPT == 1 && te - t0 < 64) ? te - t0 : 64;
    const int g = 64 / p;
    const int grp = (int)threadIdx.x / p;            // >= g: idle lanes past g * p
    const int slot = (int)threadIdx.x - grp * p;
    typename Op::Tgt tg[TPT];
    typename Op::Acc acc[TPT];
    bool act[TPT];
#pragma unroll
    for (int u = 0; u < TPT; ++u) {
        const int ti = t0 + slot + 64 * u;
        act[u] = ti < te && grp == 0;
        const int tl = ti < te ? ti : t0;
        tg[u].x = sx[tl]; tg[u].y = sy[tl]; tg[u].z = sz[tl];
        tg[u].ax = sax[tl]; tg[u].ay = say[tl]; tg[u].az = saz[tl];
        Op::init(acc[u]);
    }
    if constexpr (KERNEL == 0) gauss_stage();   // read after the first tile barrier
    const int cx = c / (ny * nz), cy = (c / nz) % ny, cz = c % nz;
    const int sl = blockIdx.z;
    for (int e = slice_off[sl]; e < slice_off[sl + 1]; ++e) {
        const int ix = cx + cols[3 * e], iy = cy + cols[3 * e + 1], dzm = cols[3 * e + 2];
        if (ix < 0 || ix >= nx || iy < 0 || iy >= ny) continue;
        const int iz0 = cz - dzm < 0 ? 0 : cz - dzm;
        const int iz1 = cz + dzm >= nz ? nz - 1 : cz + dzm;
        const int kb = (ix * ny + iy) * nz;
        const int sb = pstart[kb + iz0], se = pstart[kb + iz1 + 1];
        for (int base = sb; base < se; base += 64) {
            const int cnt = (se - base < 64) ? se - base : 64;
            __syncthreads();
            if (threadIdx.x < cnt) {
                const int j = base + threadIdx.x;
                tile[threadIdx.x] = Op::make(KERNEL, sx[j], sy[j], sz[j], sax[j], say[j],
                                             saz[j], ssg[j]);
            }
            __syncthreads();
            if (g == 1) {
                for (int k = 0; k < cnt; ++k) {
                    const PartRec sr = tile[k];
#pragma unroll
                    for (int u = 0; u < TPT; ++u) Op::pair(tg[u], sr, acc[u]);
                }
            } else {
                for (int k = grp; k < cnt; k += g) {
                    const PartRec sr = tile[k];
#pragma unroll
                    for (int u = 0; u < TPT; ++u) Op::pair(tg[u], sr, acc[u]);
                }
            }
        }
    }
    if (TPT == 1 && g > 1) {
        constexpr int kW = Op::kComp + 3;
        __shared__ double red[64 * kW];
#pragma unroll
        for (int k = 0; k < Op::kComp; ++k) red[threadIdx.x * kW + k] = acc[0].v[k];
#pragma unroll
        for (int k = 0; k < 3; ++k) red[threadIdx.x * kW + Op::kComp + k] = acc[0].g[k];
        __syncthreads();
        if (grp == 0) {
            for (int q = 1; q < g; ++q) {
                const double *r = red + (slot + q * p) * kW;
#pragma unroll
                for (int k = 0; k < Op::kComp; ++k) acc[0].v[k] += r[k];
#pragma unroll
                for (int k = 0; k < 3; ++k) acc[0].g[k] += r[Op::kComp + k];
            }
        }
    }
    double *dst = npart + (int64_t)sl * 3 * n;
#pragma unroll
    for (int u = 0; u < TPT; ++u) {
        Op::finish(tg[u], acc[u]);
        if (act[u]) {
            const int64_t ti = t0 + slot + 64 * u;
            for (int k = 0; k < 3; ++k) dst[3 * ti + k] = Op::result(acc[u], k);
        }
    }
}


// Near-field pairs of a binning: every occupied target cell's population
// times the particles of its stencil columns (what k_near evaluates), summed
// exactly in integers (one atomic add per wavefront).
__global__ void k_near_pairs(int64_t nocc, const int32_t *__restrict__ occ,
                             const int32_t *__restrict__ cstart, const int32_t *__restrict__ cend,
                             const int32_t *__restrict__ pstart, const int32_t *__restrict__ cols,
                             int ncol, int nx, int ny, int nz, unsigned long long *cnt)
{
    const int64_t o = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    unsigned long long v = 0;
    if (o < nocc) {
        const int c = occ[o];
        const int cx = c / (ny * nz), cy = (c / nz) % ny, cz = c % nz;
        long long src = 0;
        for (int e = 0; e < ncol; ++e) {
            const int ix = cx + cols[3 * e], iy = cy + cols[3 * e + 1], dzm = cols[3 * e + 2];
            if (ix < 0 || ix >= nx || iy < 0 || iy >= ny) continue;
            const int iz0 = cz - dzm < 0 ? 0 : cz - dzm;
            const int iz1 = cz + dzm >= nz ? nz - 1 : cz + dzm;
            const int kb = (ix * ny + iy) * nz;
            src += pstart[kb + iz1 + 1] - pstart[kb + iz0];
        }
        v = (unsigned long long)(cend[c] - cstart[c]) * (unsigned long long)src;
    }
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if ((threadIdx.x & 63) == 0 && v) atomicAdd(cnt, v);
}

// The pairs of one near-field pass (the binning's) into the running total.
__global__ void k_add_near_pairs(unsigned long long *cnt) { cnt[1] += cnt[0]; }

// Far field at each particle (sorted order) from the far cells of one slice of
// the occupied cells (grid.y), as that slice's partial sums: mode 0 velocity;
// 2 stretching J alpha; 3 vorticity curl u.  One particle per lane, the
// slice's cell moments staged in LDS.
template <int MODE>
__global__ __launch_bounds__(256) void k_far_part(
    int64_t n, int64_t nocc, int nfs, const int32_t *__restrict__ key_s,
    const int32_t *__restrict__ occ_of, const double *__restrict__ sx,
    const double *__restrict__ sy, const double *__restrict__ sz, const double *__restrict__ sax,
    const double *__restrict__ say, const double *__restrict__ saz, const double *__restrict__ mom,
    double rc2, double rl2, const double *__restrict__ loc, double *__restrict__ fpart)
{
    // 64-cell tiles (9.7 KB of LDS) keep ~7 wavefronts per SIMD resident
    __shared__ double tile[kFarTile * kMomTile];
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const bool active = i < n;
    const int64_t il = active ? i : n - 1;
    const int64_t t = occ_of[key_s[il]];
    const double *mt = mom + t * kMomStride;
    const double c0 = mt[16], c1 = mt[17], c2 = mt[18];
    const double x0 = sx[il], x1 = sy[il], x2 = sz[il];
    const double al[3] = {sax[il], say[il], saz[il]};
    const int64_t sb = nocc * blockIdx.y / nfs, se = nocc * (blockIdx.y + 1) / nfs;
    double r[3] = {0.0, 0.0, 0.0};
    for (int64_t s0 = sb; s0 < se; s0 += kFarTile) {
        const int cnt = (int)((se - s0) < kFarTile ? (se - s0) : kFarTile);
        __syncthreads();
        for (int e = threadIdx.x; e < cnt * kMomTile; e += 256) {
            const int k = e / kMomTile, q = e - k * kMomTile;
            tile[e] = mom[(s0 + k) * kMomStride + q];
        }
        __syncthreads();
        for (int k = 0; k < cnt; ++k) {
            const double *ms = tile + k * kMomTile;
            const double dx = c0 - ms[16], dy = c1 - ms[17], dz = c2 - ms[18];
            const double g2 = dx * dx + dy * dy + dz * dz;
            if (g2 <= rc2 || g2 > rl2) continue;
            far_term<MODE>(ms, x0, x1, x2, al, r);
        }
    }
    if (!active) return;
    if (loc && blockIdx.y == 0) {
        const double dd[3] = {x0 - c0, x1 - c1, x2 - c2};
        local_term<MODE>(loc + t * kLocStride, dd, al, r);
    }
    double *dst = fpart + (int64_t)blockIdx.y * 3 * n;
    for (int a = 0; a < 3; ++a) dst[3 * i + a] = r[a];
}

// Far field with the local expansions on: one workgroup per (target cell,
// 256-particle chunk); the cells of the target cell's band list (kappa h <
// distance <= kappa_local h) staged in LDS, each particle's sum per lane, plus
// the target cell's local expansion (everything beyond the band).
template <int MODE>
__global__ __launch_bounds__(256) void k_far_band(
    int64_t n, const int32_t *__restrict__ occ, const int32_t *__restrict__ cstart,
    const int32_t *__restrict__ cend, const int32_t *__restrict__ boff,
    const int32_t *__restrict__ bs, const double *__restrict__ sx,
    const double *__restrict__ sy, const double *__restrict__ sz, const double *__restrict__ sax,
    const double *__restrict__ say, const double *__restrict__ saz, const double *__restrict__ mom,
    const double *__restrict__ loc, double *__restrict__ fpart)
{
    __shared__ double tile[kFarTile * kMomTile];
    const int64_t t = blockIdx.x;
    const int c = occ[t];
    const int64_t p0 = cstart[c] + (int64_t)blockIdx.y * 256, p1 = cend[c];
    if (p0 >= p1) return;
    const int64_t i = p0 + threadIdx.x;
    const bool active = i < p1;
    const int64_t il = active ? i : p0;
    const double *mt = mom + t * kMomStride;
    const double x0 = sx[il], x1 = sy[il], x2 = sz[il];
    const double al[3] = {sax[il], say[il], saz[il]};
    double r[3] = {0.0, 0.0, 0.0};
    const int64_t b0 = boff[t], b1 = boff[t + 1];
    for (int64_t s0 = b0; s0 < b1; s0 += kFarTile) {
        const int cnt = (int)((b1 - s0) < kFarTile ? (b1 - s0) : kFarTile);
        __syncthreads();
        for (int e = threadIdx.x; e < cnt * kMomTile; e += 256) {
            const int k = e / kMomTile, q = e - k * kMomTile;
            tile[e] = mom[(int64_t)bs[s0 + k] * kMomStride + q];
        }
        __syncthreads();
        for (int k = 0; k < cnt; ++k) far_term<MODE>(tile + k * kMomTile, x0, x1, x2, al, r);
    }
    if (!active) return;
    const double dd[3] = {x0 - mt[16], x1 - mt[17], x2 - mt[18]};
    local_term<MODE>(loc + t * kLocStride, dd, al, r);
    for (int a = 0; a < 3; ++a) fpart[3 * i + a] = r[a];
}

// Near slices + far slices (+ the learned residual of the particle's cell) in
// a fixed order, scattered to the original particle order.
__global__ void k_combine(int64_t n, int nns, int nfs, const int32_t *__restrict__ key_s,
                          const int32_t *__restrict__ occ_of, const int32_t *__restrict__ perm,
                          const double *__restrict__ npart, const double *__restrict__ fpart,
                          const double *__restrict__ res, double *__restrict__ out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double r[3] = {0.0, 0.0, 0.0};
    for (int sl = 0; sl < nns; ++sl)
        for (int a = 0; a < 3; ++a) r[a] += npart[(int64_t)sl * 3 * n + 3 * i + a];
    for (int sl = 0; sl < nfs; ++sl)
        for (int a = 0; a < 3; ++a) r[a] += fpart[(int64_t)sl * 3 * n + 3 * i + a];
    if (res) {
        const int64_t t = occ_of[key_s[i]];
        for (int a = 0; a < 3; ++a) r[a] += res[3 * t + a];
    }
    const int64_t j = perm[i];
    for (int a = 0; a < 3; ++a) out[3 * j + a] = r[a];
}

__device__ __forceinline__ uint64_t mix64(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// Training samples: a random shell pair and a random particle of its target
// cell; the exact velocity of the source cell's particles at that particle
// (singular kernel, direct sum, one wavefront per sample) minus the cell's
// expansion, divided by the residual scale.  Unusable samples (zero scale,
// non-finite or |target| > 1e4) are zeroed.
__global__ __launch_bounds__(64) void k_train_samples(
    int64_t nlist, const int32_t *__restrict__ lt, const int32_t *__restrict__ ls,
    const int32_t *__restrict__ occ, const int32_t *__restrict__ cstart,
    const int32_t *__restrict__ cend, const double *__restrict__ sx, const double *__restrict__ sy,
    const double *__restrict__ sz, const double *__restrict__ sax, const double *__restrict__ say,
    const double *__restrict__ saz, const double *__restrict__ mom, double h, uint64_t seed,
    float *__restrict__ X, float *__restrict__ T)
{
    const int64_t k = blockIdx.x;
    __shared__ int64_t pick[2];
    if (threadIdx.x == 0) {
        const int64_t e = (int64_t)(mix64(seed ^ (2 * k + 1)) % (uint64_t)nlist);
        const int ct = occ[lt[e]];
        const int pop = cend[ct] - cstart[ct];
        pick[0] = cstart[ct] + (int64_t)(mix64(seed ^ (2 * k + 2)) % (uint64_t)pop);
        pick[1] = ls[e];
    }
    __syncthreads();
    const int64_t p = pick[0], s = pick[1];
    const double x0 = sx[p], x1 = sy[p], x2 = sz[p];
    const int c = occ[s];
    double u[3] = {0, 0, 0};
    for (int j = cstart[c] + threadIdx.x; j < cend[c]; j += 64) {
        const double d[3] = {x0 - sx[j], x1 - sy[j], x2 - sz[j]};
        const double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
        const double ir = 1.0 / sqrt(r2);
        const double k3 = -kInv4Pi * ir * ir * ir;
        u[0] += k3 * (d[1] * saz[j] - d[2] * say[j]);
        u[1] += k3 * (d[2] * sax[j] - d[0] * saz[j]);
        u[2] += k3 * (d[0] * say[j] - d[1] * sax[j]);
    }
    for (int a = 0; a < 3; ++a)
        for (int off = 32; off > 0; off >>= 1) u[a] += __shfl_xor(u[a], off, 64);
    if (threadIdx.x == 0) {
        const double *ms = mom + s * kMomStride;
        float f[kFeat];
        const double sc = far_features(ms, x0, x1, x2, h, f);
        double ua[3] = {0.0, 0.0, 0.0};
        far_expansion(ms, x0, x1, x2, ua, nullptr);
        double tv[3];
        bool ok = sc > 0.0;
        for (int a = 0; a < 3; ++a) {
            tv[a] = ok ? (u[a] - ua[a]) / sc : 0.0;
            ok = ok && isfinite(tv[a]) && fabs(tv[a]) <= 1e4;
        }
        for (int a = 0; a < kFeat; ++a) X[k * kFeat + a] = ok ? f[a] : 0.0f;
        for (int a = 0; a < kOutW; ++a) T[k * kOutW + a] = (ok && a < 3) ? (float)tv[a] : 0.0f;
    }
}

// max |a - b| and max |b| over n x 3 (one workgroup, fixed order); a
// non-finite a makes the first +inf.
__global__ __launch_bounds__(256) void k_maxdiff(int64_t n, const double *__restrict__ a,
                                                 const double *__restrict__ b,
                                                 double *__restrict__ out)
{
    __shared__ double s0[256], s1[256];
    double m0 = 0.0, m1 = 0.0;
    for (int64_t i = threadIdx.x; i < 3 * n; i += 256) {
        const double d = fabs(a[i] - b[i]);
        m0 = isfinite(d) ? fmax(m0, d) : INFINITY;
        m1 = fmax(m1, fabs(b[i]));
    }
    s0[threadIdx.x] = m0;
    s1[threadIdx.x] = m1;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            s0[threadIdx.x] = fmax(s0[threadIdx.x], s0[threadIdx.x + w]);
            s1[threadIdx.x] = fmax(s1[threadIdx.x], s1[threadIdx.x + w]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = s0[0];
        out[1] = s1[0];
    }
}

// ------------------------------------------------------------- host side

static int grow(fv_ctx *ctx, void **p, size_t *cap, size_t bytes)
{
    if (*cap >= bytes) return FV_OK;
    if (*p) {
        FV_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
        FV_HIP_CHECK(ctx, hipFree(*p));
        *p = nullptr;
        *cap = 0;
    }
    bytes = bytes + bytes / 4 + 256;
    FV_HIP_CHECK(ctx, hipMalloc(p, bytes));
    *cap = bytes;
    return FV_OK;
}

// Near stencil as columns: for every (dx, dy) the cells dz in [-dz_max, dz_max]
// with dx^2 + dy^2 + dz^2 <= kappa^2 (grid-centre distance <= kappa h), each
// axis clamped to the grid (offsets beyond it can never land on a cell);
// grouped into at most kMaxSlices contiguous dx ranges.
static int update_stencil(fv_closure *cl, fv_ctx *ctx)
{
    const int K = (int)ceil(cl->kappa);
    const int kx = K < cl->nx - 1 ? K : cl->nx - 1, ky = K < cl->ny - 1 ? K : cl->ny - 1,
              kz = K < cl->nz - 1 ? K : cl->nz - 1;
    if (cl->dstencil && kx == cl->skx && ky == cl->sky && kz == cl->skz) return FV_OK;
    cl->stencil.clear();
    cl->slice_off.clear();
    const double k2 = cl->kappa * cl->kappa + 1e-9;
    const int ndx = 2 * kx + 1;
    cl->nsl = ndx < kMaxSlices ? ndx : kMaxSlices;
    for (int sl = 0; sl < cl->nsl; ++sl) {
        cl->slice_off.push_back((int)(cl->stencil.size() / 3));
        const int d0 = -kx + (int)((int64_t)ndx * sl / cl->nsl);
        const int d1 = -kx + (int)((int64_t)ndx * (sl + 1) / cl->nsl);
        for (int i = d0; i < d1; ++i)
            for (int j = -ky; j <= ky; ++j) {
                const double r2 = k2 - (double)i * i - (double)j * j;
                if (r2 < 0.0) continue;
                int dz = (int)floor(sqrt(r2));
                while ((double)(dz + 1) * (dz + 1) <= r2) ++dz;
                while (dz > 0 && (double)dz * dz > r2) --dz;
                if (dz > kz) dz = kz;
                cl->stencil.push_back(i);
                cl->stencil.push_back(j);
                cl->stencil.push_back(dz);
            }
    }
    cl->slice_off.push_back((int)(cl->stencil.size() / 3));
    if (cl->dstencil) {
        FV_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
        FV_HIP_CHECK(ctx, hipFree(cl->dstencil));
        cl->dstencil = nullptr;
    }
    const size_t nst = cl->stencil.size(), nso = cl->slice_off.size();
    FV_HIP_CHECK(ctx, hipMalloc(&cl->dstencil, sizeof(int32_t) * (nst + nso)));
    FV_HIP_CHECK(ctx, hipMemcpyAsync(cl->dstencil, cl->stencil.data(), sizeof(int32_t) * nst,
                                     hipMemcpyHostToDevice, ctx->stream));
    FV_HIP_CHECK(ctx, hipMemcpyAsync(cl->dstencil + nst, cl->slice_off.data(),
                                     sizeof(int32_t) * nso, hipMemcpyHostToDevice, ctx->stream));
    FV_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    cl->skx = kx;
    cl->sky = ky;
    cl->skz = kz;
    return FV_OK;
}

static int closure_residual(fv_vpm *f, bool use_mlp);

static int build_local(fv_closure *cl, fv_ctx *ctx);

// The field's stream waits for the pending online training (closure_train)
// before anything reads the weights it writes or rewrites the binning it reads.
static int join_training(fv_closure *cl, fv_ctx *ctx)
{
    if (!cl->train_pending) return FV_OK;
    FV_HIP_CHECK(ctx, hipStreamWaitEvent(ctx->stream, cl->ev_train, 0));
    cl->train_pending = false;
    return FV_OK;
}

// Bins the particles (cells, moments, shell pair list) for the current state
// and, when `with_res`, computes the learned residual.  With `stage` (the RK
// stages of one advect step, whose particles move a small fraction of a cell)
// the step's binning is kept: cell membership, the shell list and the learned
// residual stay, positions, strengths and moments are refreshed.
//
// A full binning is sized on the device: the bounding box, the grid, the cell
// keys, the stable sort, the cell ranges, the occupied cells, their moments
// and the shell / band counts run with launches and buffers sized by bounds
// (the particle launch count, the grid's cell limit), every kernel taking the
// exact sizes from the device grid record; the host reads that record once,
// after the last count, and sizes the pair lists and the later passes by it.
static int closure_build(fv_vpm *f, bool use_mlp, bool with_res = true, bool stage = false)
{
    fv_closure *cl = f->closure;
    fv_ctx *ctx = f->ctx;
    if (cl->built_epoch == f->state_epoch) return with_res ? closure_residual(f, use_mlp) : FV_OK;
    int rc = join_training(cl, ctx);
    if (rc) return rc;
    if (stage && cl->binned_step == cl->steps && cl->nocc > 0) {
        rc = f->np_exact ? FV_OK : vpm_sync_count(f);
        if (rc) return rc;
        const int64_t n = f->np_host;
        if (cl->n == n && n > 0) {
            const unsigned gb = (unsigned)ceil_div(n, 256);
            hipLaunchKernelGGL(k_gather, dim3(gb), dim3(256), 0, ctx->stream, n, cl->grid, cl->perm,
                               f->x, f->y, f->z, f->gx, f->gy, f->gz, f->sig, cl->sx, cl->sy,
                               cl->sz, cl->sax, cl->say, cl->saz, cl->ssg);
            FV_LAUNCH_CHECK(ctx);
            hipLaunchKernelGGL(k_moments, dim3((unsigned)cl->nocc), dim3(64), 0, ctx->stream,
                               cl->grid, cl->occ, cl->cstart, cl->cend, cl->sx, cl->sy, cl->sz,
                               cl->sax, cl->say, cl->saz, cl->mom);
            FV_LAUNCH_CHECK(ctx);
            rc = build_local(cl, ctx);
            if (rc) return rc;
            cl->built_epoch = f->state_epoch;
            return with_res ? closure_residual(f, use_mlp) : FV_OK;
        }
    }
    const int64_t nu = launch_count(f);   // >= the device particle count
    cl->bin_id++;
    cl->binned_step = cl->steps;
    if (nu == 0) {
        cl->n = 0;
        cl->nocc = 0;
        cl->nlist = 0;
        cl->nband = 0;
        cl->built_epoch = f->state_epoch;
        return FV_OK;
    }
    // bounding box and grid
    const int nbb = 64;
    rc = grow(ctx, &cl->tmp, &cl->tmp_bytes, 4096);
    if (rc) return rc;
    if (!cl->bbox) {
        FV_HIP_CHECK(ctx, hipMalloc(&cl->bbox, sizeof(double) * (6 * nbb + 8) + sizeof(ClGrid)));
        cl->near_cnt = reinterpret_cast<unsigned long long *>(cl->bbox + 6 * nbb);
        cl->grid = reinterpret_cast<ClGrid *>(cl->bbox + 6 * nbb + 8);
        FV_HIP_CHECK(ctx, hipMemsetAsync(cl->near_cnt, 0, 2 * sizeof(unsigned long long),
                                         ctx->stream));
    }
    hipLaunchKernelGGL(k_bbox_partial, dim3(nbb), dim3(256), 0, ctx->stream, f->np_dev, f->x,
                       f->y, f->z, cl->bbox);
    FV_LAUNCH_CHECK(ctx);
    hipLaunchKernelGGL(k_grid, dim3(1), dim3(64), 0, ctx->stream, cl->bbox, nbb, f->np_dev, cl->h,
                       (double)cl->max_cells, cl->grid, cl->near_cnt);
    FV_LAUNCH_CHECK(ctx);
    // bounds: cells of a grid, occupied cells
    const int64_t ncb = 8 * cl->max_cells < 262144 ? 8 * cl->max_cells : 262144;
    const int64_t nob = nu < ncb ? nu : ncb;
    // particle-side buffers
    const size_t need = (size_t)nu * (4 * sizeof(int32_t) + (7 + 6 * kMaxSlices) * sizeof(double)) +
                        1024;
    rc = grow(ctx, &cl->block, &cl->block_bytes, need);
    if (rc) return rc;
    char *p = static_cast<char *>(cl->block);
    auto takeD = [&](size_t k) { double *r = reinterpret_cast<double *>(p); p += k * sizeof(double); return r; };
    cl->sx = takeD(nu); cl->sy = takeD(nu); cl->sz = takeD(nu);
    cl->sax = takeD(nu); cl->say = takeD(nu); cl->saz = takeD(nu); cl->ssg = takeD(nu);
    cl->npart = takeD(3 * kMaxSlices * nu);
    cl->fpart = takeD(3 * kMaxSlices * nu);
    cl->key = reinterpret_cast<int32_t *>(p); p += nu * sizeof(int32_t);
    cl->key_s = reinterpret_cast<int32_t *>(p); p += nu * sizeof(int32_t);
    cl->idx = reinterpret_cast<int32_t *>(p); p += nu * sizeof(int32_t);
    cl->perm = reinterpret_cast<int32_t *>(p); p += nu * sizeof(int32_t);
    const unsigned gb = (unsigned)ceil_div(nu, 256);
    hipLaunchKernelGGL(k_cell_keys, dim3(gb), dim3(256), 0, ctx->stream, nu, cl->grid, f->x, f->y,
                       f->z, cl->key, cl->idx);
    FV_LAUNCH_CHECK(ctx);
    // stable radix sort of (cell, index); the slots past the count sort last
    size_t tb = 0;
    FV_HIP_CHECK(ctx, hipcub::DeviceRadixSort::SortPairs(nullptr, tb, cl->key, cl->key_s, cl->idx,
                                                        cl->perm, (int)nu, 0, kKeyBits,
                                                        ctx->stream));
    rc = grow(ctx, &cl->tmp, &cl->tmp_bytes, tb);
    if (rc) return rc;
    FV_HIP_CHECK(ctx, hipcub::DeviceRadixSort::SortPairs(cl->tmp, tb, cl->key, cl->key_s, cl->idx,
                                                        cl->perm, (int)nu, 0, kKeyBits,
                                                        ctx->stream));
    // per-cell arrays (for the grid's cell limit): cell starts, occupied cells
    const size_t cneed = ((size_t)ncb * 3 + 2) * sizeof(int32_t) + 1024;
    rc = grow(ctx, &cl->cells, &cl->cells_bytes, cneed);
    if (rc) return rc;
    cl->pstart = static_cast<int32_t *>(cl->cells);
    cl->cstart = cl->pstart;
    cl->cend = cl->pstart + 1;
    cl->occ = cl->pstart + ncb + 1;
    cl->occ_of = cl->occ + ncb;
    hipLaunchKernelGGL(k_cell_bounds, dim3((unsigned)ceil_div(ncb + 1, 256)), dim3(256), 0,
                       ctx->stream, ncb, cl->grid, cl->key_s, cl->pstart);
    FV_LAUNCH_CHECK(ctx);
    hipLaunchKernelGGL(k_gather, dim3(gb), dim3(256), 0, ctx->stream, nu, cl->grid, cl->perm, f->x,
                       f->y, f->z, f->gx, f->gy, f->gz, f->sig, cl->sx, cl->sy, cl->sz, cl->sax,
                       cl->say, cl->saz, cl->ssg);
    FV_LAUNCH_CHECK(ctx);
    hipLaunchKernelGGL(k_occupied, dim3(1), dim3(1024), 0, ctx->stream, cl->grid, cl->cstart,
                       cl->cend, cl->occ, cl->occ_of);
    FV_LAUNCH_CHECK(ctx);
    // moments, residual and local expansions per cell, shell and band counts /
    // offsets (for the occupied-cell bound)
    const size_t fneed = (size_t)nob * (kMomStride + 3 + kLocStride) * sizeof(double) +
                         (size_t)4 * (nob + 1) * sizeof(int32_t) + 1024;
    rc = grow(ctx, &cl->far, &cl->far_bytes, fneed);
    if (rc) return rc;
    cl->mom = static_cast<double *>(cl->far);
    cl->res = cl->mom + nob * kMomStride;
    cl->loc = cl->res + 3 * nob;
    cl->lcnt = reinterpret_cast<int32_t *>(cl->loc + kLocStride * nob);
    cl->loff = cl->lcnt + (nob + 1);
    cl->bcnt = cl->loff + (nob + 1);
    cl->boff = cl->bcnt + (nob + 1);
    hipLaunchKernelGGL(k_moments, dim3((unsigned)nob), dim3(64), 0, ctx->stream, cl->grid, cl->occ,
                       cl->cstart, cl->cend, cl->sx, cl->sy, cl->sz, cl->sax, cl->say, cl->saz,
                       cl->mom);
    FV_LAUNCH_CHECK(ctx);
    const bool band = cl->kappa_local > 0.0;
    hipLaunchKernelGGL(k_shell_count, dim3((unsigned)(nob + 1)), dim3(64), 0, ctx->stream,
                       cl->grid, cl->mom, cl->kappa, cl->kappa_local, cl->lcnt,
                       band ? cl->bcnt : (int32_t *)nullptr);
    FV_LAUNCH_CHECK(ctx);
    hipLaunchKernelGGL(k_scan_lists, dim3(1), dim3(1024), 0, ctx->stream, cl->grid, cl->lcnt,
                       cl->loff, band ? cl->bcnt : (const int32_t *)nullptr, cl->boff);
    FV_LAUNCH_CHECK(ctx);
    // the binning's one host read
    ClGrid G;
    FV_HIP_CHECK(ctx, hipMemcpyAsync(&G, cl->grid, sizeof G, hipMemcpyDeviceToHost, ctx->stream));
    FV_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    if (G.bad) return set_error(ctx, FV_E_ARG, "closure: non-finite particle positions");
    f->np_host = G.n;
    f->np_upper = G.n;
    f->np_exact = true;
    const int64_t n = G.n;
    cl->n = n;
    cl->hh = G.h;
    cl->ox = G.ox;
    cl->oy = G.oy;
    cl->oz = G.oz;
    cl->nx = G.nx;
    cl->ny = G.ny;
    cl->nz = G.nz;
    cl->ncell = G.ncell;
    cl->nocc = G.nocc;
    cl->maxpop = G.maxpop;
    cl->nlist = G.nl;
    cl->nband = band ? G.nb : 0;
    if (n == 0) {
        cl->nocc = 0;
        cl->nlist = 0;
        cl->nband = 0;
        cl->built_epoch = f->state_epoch;
        return FV_OK;
    }
    const int64_t nocc = cl->nocc;
    const double h = cl->hh;
    rc = build_local(cl, ctx);
    if (rc) return rc;
    // shell pair list (geometry only) and far band list (source cells with
    // kappa h < distance <= kappa_local h), in one pass
    const double rc2 = cut2(cl->kappa, h);
    const int64_t nl = cl->nlist;
    if (nl > 0) {
        const size_t lneed = (size_t)nl * (2 * sizeof(int32_t) + sizeof(double) +
                                           (kFeat + kOutW) * sizeof(float)) + 1024;
        rc = grow(ctx, &cl->list, &cl->list_bytes, lneed);
        if (rc) return rc;
        cl->lk = static_cast<double *>(cl->list);
        cl->X = reinterpret_cast<float *>(cl->lk + nl);
        cl->Y = cl->X + (size_t)nl * kFeat;
        cl->lt = reinterpret_cast<int32_t *>(cl->Y + (size_t)nl * kOutW);
        cl->ls = cl->lt + nl;
    }
    const int64_t nb = cl->nband;
    if (nb > 0) {
        rc = grow(ctx, &cl->band, &cl->band_bytes, (size_t)2 * (nb + 1) * sizeof(int32_t));
        if (rc) return rc;
        cl->bt = static_cast<int32_t *>(cl->band);
        cl->bs = cl->bt + (nb + 1);
    }
    if (nl > 0 || nb > 0) {
        hipLaunchKernelGGL(k_shell_fill, dim3((unsigned)nocc), dim3(64), 0, ctx->stream, nocc,
                           cl->mom, rc2, rc2 * kShell * kShell,
                           nb > 0 ? cut2(cl->kappa_local, h) : 0.0, cl->loff,
                           nl > 0 ? cl->lt : (int32_t *)nullptr, cl->ls, cl->boff,
                           nb > 0 ? cl->bt : (int32_t *)nullptr, cl->bs);
        FV_LAUNCH_CHECK(ctx);
    }
    rc = update_stencil(cl, ctx);
    if (rc) return rc;
    // near-field pairs of this binning (every pass of it evaluates them; the
    // count was cleared by k_grid)
    hipLaunchKernelGGL(k_near_pairs, dim3((unsigned)ceil_div(nocc, 256)), dim3(256), 0,
                       ctx->stream, nocc, cl->occ, cl->cstart, cl->cend, cl->pstart,
                       cl->dstencil, (int)(cl->stencil.size() / 3), cl->nx, cl->ny, cl->nz,
                       cl->near_cnt);
    FV_LAUNCH_CHECK(ctx);
    cl->built_epoch = f->state_epoch;
    return with_res ? closure_residual(f, use_mlp) : FV_OK;
}

// Learned residual per occupied cell (zero without a network or shell pairs),
// once per binning and network version.
static int closure_residual(fv_vpm *f, bool use_mlp)
{
    fv_closure *cl = f->closure;
    fv_ctx *ctx = f->ctx;
    use_mlp = use_mlp && cl->mlp && cl->nlist > 0;
    if (cl->res_bin == cl->bin_id && cl->res_mlp == (int)use_mlp &&
        (!use_mlp || cl->res_version == cl->mlp_version))
        return FV_OK;
    const int64_t nocc = cl->nocc;
    if (nocc == 0) return FV_OK;
    int rc = join_training(cl, ctx);
    if (rc) return rc;
    if (use_mlp) {
        TimedLaunch tl(ctx, "closure_residual");
        const int64_t nl = cl->nlist;
        hipLaunchKernelGGL(k_shell_inputs, dim3((unsigned)ceil_div(nl, 256)), dim3(256), 0,
                           ctx->stream, nl, cl->lt, cl->ls, cl->mom, cl->hh, cl->X, cl->lk);
        FV_LAUNCH_CHECK(ctx);
        {
            MlpOn on(cl->mlp, ctx->stream);
            rc = fv_mlp_forward_dev(cl->mlp, cl->X, nl, cl->Y);
        }
        if (rc) return rc;
        hipLaunchKernelGGL(k_shell_reduce, dim3((unsigned)nocc), dim3(64), 0, ctx->stream, nocc,
                           cl->loff, cl->lk, cl->Y, cl->res);
        FV_LAUNCH_CHECK(ctx);
        tl.done();
        cl->far_evals += (double)nl;
    } else {
        rc = fill(ctx, cl->res, 3 * nocc, 0.0);
        if (rc) return rc;
    }
    cl->res_bin = cl->bin_id;
    cl->res_mlp = (int)use_mlp;
    cl->res_version = cl->mlp_version;
    return FV_OK;
}

template <int K, int P, int M>
static void launch_near(fv_closure *cl, fv_ctx *ctx, int tpt)
{
    const size_t nst = cl->stencil.size();
    const unsigned chunks = (unsigned)ceil_div(cl->maxpop > 0 ? cl->maxpop : 1, 64 * tpt);
    const dim3 grid((unsigned)cl->nocc, chunks, (unsigned)cl->nsl);
#define FV_NEAR_T(T)                                                                             \
    hipLaunchKernelGGL((k_near<K, P, M, T>), grid, dim3(64), 0, ctx->stream, cl->n, cl->nocc,   \
                       cl->occ, cl->cstart, cl->cend, cl->pstart, cl->dstencil,                 \
                       cl->dstencil + nst, cl->nx, cl->ny, cl->nz, cl->sx, cl->sy, cl->sz,     \
                       cl->sax, cl->say, cl->saz, cl->ssg, cl->npart)
    if (tpt == 4) FV_NEAR_T(4);
    else if (tpt == 1) FV_NEAR_T(1);
    else FV_NEAR_T(2);
#undef FV_NEAR_T
}

// Far slices: enough source-cell ranges for ~8 wavefronts per SIMD.
static int far_slices(const fv_closure *cl)
{
    int64_t fs = ceil_div(int64_t(1) << 19, cl->n > 0 ? cl->n : 1);
    if (fs > kMaxSlices) fs = kMaxSlices;
    if (fs > cl->nocc) fs = cl->nocc;
    return fs < 1 ? 1 : (int)fs;
}

template <int M>
static void launch_far(fv_closure *cl, fv_ctx *ctx, double rc2, int nfs)
{
    if (cl->kappa_local > 0.0) {
        const unsigned chunks = (unsigned)ceil_div(cl->maxpop > 0 ? cl->maxpop : 1, 256);
        hipLaunchKernelGGL((k_far_band<M>), dim3((unsigned)cl->nocc, chunks), dim3(256), 0,
                           ctx->stream, cl->n, cl->occ, cl->cstart, cl->cend, cl->boff, cl->bs,
                           cl->sx, cl->sy, cl->sz, cl->sax, cl->say, cl->saz, cl->mom, cl->loc,
                           cl->fpart);
        return;
    }
    const bool lx = cl->kappa_local > 0.0;
    const double rl = cl->kappa_local * cl->hh;
    const double rl2 = lx ? rl * rl * (1.0 + 1e-12) : INFINITY;
    hipLaunchKernelGGL((k_far_part<M>), dim3((unsigned)ceil_div(cl->n, 256), (unsigned)nfs),
                       dim3(256), 0, ctx->stream, cl->n, cl->nocc, nfs, cl->key_s, cl->occ_of,
                       cl->sx, cl->sy, cl->sz, cl->sax, cl->say, cl->saz, cl->mom, rc2, rl2,
                       lx ? cl->loc : (const double *)nullptr, cl->fpart);
}

// Local expansions of every occupied cell (after its moments are current).
static int build_local(fv_closure *cl, fv_ctx *ctx)
{
    if (!(cl->kappa_local > 0.0) || cl->nocc == 0) return FV_OK;
    const double rl = cl->kappa_local * cl->hh;
    const int64_t nocc = cl->nocc;
    // about 1024 workgroups of four wavefronts: target blocks of 64 cells x
    // source chunks
    const int64_t tb = ceil_div(nocc, 64);
    int64_t nchunk = ceil_div(1024, tb);
    if (nchunk > 64) nchunk = 64;
    if (nchunk > nocc) nchunk = nocc;
    const int64_t chunk = ceil_div(nocc, nchunk);
    nchunk = ceil_div(nocc, chunk);
    int rc = grow(ctx, &cl->lpart, &cl->lpart_bytes, sizeof(double) * 30 * (size_t)(nchunk * nocc));
    if (rc) return rc;
    double *lp = static_cast<double *>(cl->lpart);
    hipLaunchKernelGGL(k_local_part, dim3((unsigned)tb, (unsigned)nchunk), dim3(256), 0, ctx->stream,
                       nocc, chunk, cl->mom, rl * rl * (1.0 + 1e-12), lp);
    FV_LAUNCH_CHECK(ctx);
    hipLaunchKernelGGL(k_local_sum, dim3((unsigned)ceil_div(nocc * kLocStride, 256)), dim3(256), 0,
                       ctx->stream, nocc, (int)nchunk, lp, cl->loc);
    FV_LAUNCH_CHECK(ctx);
    return FV_OK;
}

// Closure sum at the particles (original order) for mode 0 (velocity), 2
// (stretching) or 3 (vorticity); kernel selects the near-field smoothing.
static int closure_pass(fv_vpm *f, int mode, int kernel, double *out, bool use_mlp,
                        bool stage = false)
{
    fv_closure *cl = f->closure;
    fv_ctx *ctx = f->ctx;
    FV_REQUIRE(ctx, mode != 4 || kernel == FV_KERNEL_GAUSSIAN_ERF,
               "the clamped-pair vorticity (mode 4) is the Gaussian field's");
    // the learned residual (velocity only) is formed after the near and far
    // passes, so the online training of this binning overlaps them
    int rc = closure_build(f, use_mlp, false, stage);
    if (rc) return rc;
    const int64_t n = cl->n;
    if (n == 0) return FV_OK;
    const int prec = f->precision;
    auto near = [&](int tpt) {
#define FV_NEAR(K, P)                                                   \
    do {                                                                \
        if (mode == 0) launch_near<K, P, 0>(cl, ctx, tpt);              \
        else if (mode == 2) launch_near<K, P, 2>(cl, ctx, tpt);         \
        else launch_near<K, P, 3>(cl, ctx, tpt);                        \
    } while (0)
        if (kernel == FV_KERNEL_WINCKELMANS) {
            if (prec == FV_PREC_FP64) FV_NEAR(1, 0); else FV_NEAR(1, 1);
        } else if (mode == 4) {
            if (prec == FV_PREC_FP64) launch_near<0, 0, 4>(cl, ctx, tpt);
            else launch_near<0, 1, 4>(cl, ctx, tpt);
        } else {
            if (prec == FV_PREC_FP64) FV_NEAR(0, 0); else FV_NEAR(0, 1);
        }
#undef FV_NEAR
    };
    // tuning slot per (mode, kernel, precision); mode 4 shares mode 3's work
    int &tsel = cl->near_tpt[mode == 4 ? 3 : mode & 3][kernel == FV_KERNEL_WINCKELMANS]
                            [prec == FV_PREC_MIXED];
    if (tsel == 0) {
        tsel = 2;
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        if (ctx->autotune && hipStreamIsCapturing(ctx->stream, &cs) == hipSuccess &&
            cs == hipStreamCaptureStatusNone) {
            // time 1, 2 and 4 targets per lane once (each twice, the second counted)
            hipEvent_t e0 = nullptr, e1 = nullptr;
            FV_HIP_CHECK(ctx, hipEventCreate(&e0));
            FV_HIP_CHECK(ctx, hipEventCreate(&e1));
            float best = INFINITY;
            for (int t : {1, 2, 4}) {
                float ms = 0.0f;
                for (int rep = 0; rep < 2; ++rep) {
                    FV_HIP_CHECK(ctx, hipEventRecord(e0, ctx->stream));
                    near(t);
                    FV_LAUNCH_CHECK(ctx);
                    FV_HIP_CHECK(ctx, hipEventRecord(e1, ctx->stream));
                    FV_HIP_CHECK(ctx, hipEventSynchronize(e1));
                    FV_HIP_CHECK(ctx, hipEventElapsedTime(&ms, e0, e1));
                }
                if (ms < best) {
                    best = ms;
                    tsel = t;
                }
            }
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
        }
    }
    {
        TimedLaunch tl(ctx, "closure_near");
        near(tsel);
        FV_LAUNCH_CHECK(ctx);
        tl.done();
    }
    hipLaunchKernelGGL(k_add_near_pairs, dim3(1), dim3(1), 0, ctx->stream, cl->near_cnt);
    FV_LAUNCH_CHECK(ctx);
    const int nfs = cl->kappa_local > 0.0 ? 1 : far_slices(cl);
    {
        TimedLaunch tl(ctx, "closure_far");
        const double rc2 = (cl->kappa * cl->hh) * (cl->kappa * cl->hh) * (1.0 + 1e-12);
        if (mode == 0) launch_far<0>(cl, ctx, rc2, nfs);
        else if (mode == 2) launch_far<2>(cl, ctx, rc2, nfs);
        else launch_far<3>(cl, ctx, rc2, nfs);
        FV_LAUNCH_CHECK(ctx);
        tl.done();
    }
    if (mode == 0) {
        rc = closure_residual(f, use_mlp);
        if (rc) return rc;
    }
    hipLaunchKernelGGL(k_combine, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, ctx->stream, n,
                       cl->nsl, nfs, cl->key_s, cl->occ_of, cl->perm, cl->npart, cl->fpart,
                       mode == 0 ? cl->res : (const double *)nullptr, out);
    FV_LAUNCH_CHECK(ctx);
    return FV_OK;
}

static int closure_train(fv_vpm *f)
{
    fv_closure *cl = f->closure;
    fv_ctx *ctx = f->ctx;
    if (!cl->mlp || cl->train_iters <= 0 || cl->n == 0 || cl->nlist == 0) return FV_OK;
    const int S = cl->train_samples;
    if (!cl->tX) {
        FV_HIP_CHECK(ctx, hipMalloc(&cl->tX, sizeof(float) * ((kFeat + kOutW) * (size_t)S + kOutW)));
        cl->tT = cl->tX + kFeat * (size_t)S;
        cl->tw = cl->tT + kOutW * (size_t)S;
        float w[kOutW] = {1, 1, 1};
        FV_HIP_CHECK(ctx, hipMemcpyAsync(cl->tw, w, sizeof w, hipMemcpyHostToDevice, ctx->stream));
        FV_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    }
    if (!cl->side) {
        FV_HIP_CHECK(ctx, hipStreamCreateWithFlags(&cl->side, hipStreamNonBlocking));
        FV_HIP_CHECK(ctx, hipEventCreateWithFlags(&cl->ev_bin, hipEventDisableTiming));
        FV_HIP_CHECK(ctx, hipEventCreateWithFlags(&cl->ev_train, hipEventDisableTiming));
    }
    // after the binning, on the side stream; the weights are next read by the
    // residual of the step's first velocity pass (join_training)
    FV_HIP_CHECK(ctx, hipEventRecord(cl->ev_bin, ctx->stream));
    FV_HIP_CHECK(ctx, hipStreamWaitEvent(cl->side, cl->ev_bin, 0));
    TimedLaunch tl(ctx, "closure_train", cl->side);
    cl->seed = cl->seed * 6364136223846793005ull + 1442695040888963407ull;
    hipLaunchKernelGGL(k_train_samples, dim3((unsigned)S), dim3(64), 0, cl->side, cl->nlist,
                       cl->lt, cl->ls, cl->occ, cl->cstart, cl->cend, cl->sx, cl->sy, cl->sz,
                       cl->sax, cl->say, cl->saz, cl->mom, cl->hh, cl->seed, cl->tX, cl->tT);
    FV_LAUNCH_CHECK(ctx);
    {
        MlpOn on(cl->mlp, cl->side);
        for (int it = 0; it < cl->train_iters; ++it) {
            int rc = fv_mlp_train_step_dev(cl->mlp, cl->tX, cl->tT, cl->tw, S, (float)cl->lr);
            if (rc) return rc;
        }
    }
    tl.done();
    FV_HIP_CHECK(ctx, hipEventRecord(cl->ev_train, cl->side));
    cl->train_pending = true;
    cl->mlp_version++;
    return FV_OK;
}

// Max-norm error of the closure velocity relative to the exact sum's max norm.
static int closure_check(fv_vpm *f, bool use_mlp, double *err)
{
    fv_closure *cl = f->closure;
    fv_ctx *ctx = f->ctx;
    int rc = closure_build(f, use_mlp);
    if (rc) return rc;
    const int64_t n = cl->n;
    *err = 0.0;
    if (n == 0) return FV_OK;
    if (!cl->vdir) {
        FV_HIP_CHECK(ctx, hipMalloc(&cl->vdir, sizeof(double) * (6 * (size_t)f->cap + 8)));
        cl->vclo = cl->vdir + 3 * (size_t)f->cap;
        cl->verr = cl->vclo + 3 * (size_t)f->cap;
    }
    rc = closure_pass(f, 0, f->kernel, cl->vclo, use_mlp);
    if (rc) return rc;
    PartSrcArgs s;
    s.x = f->x; s.y = f->y; s.z = f->z;
    s.ax = f->gx; s.ay = f->gy; s.az = f->gz;
    s.sig = f->sig; s.stride = 1; s.kernel = f->kernel;
    PartTgtArgs t;
    t.x = f->x; t.y = f->y; t.z = f->z;
    t.ax = f->gx; t.ay = f->gy; t.az = f->gz; t.stride = 1;
    rc = part_sum(ctx, f->kernel, f->precision, 0, s, t, n, n, nullptr, nullptr, cl->vdir, S_TMP1,
                  "closure_check");
    if (rc) return rc;
    f->particle_pairs += (double)n * (double)n;
    hipLaunchKernelGGL(k_maxdiff, dim3(1), dim3(256), 0, ctx->stream, n, cl->vclo, cl->vdir,
                       cl->verr);
    FV_LAUNCH_CHECK(ctx);
    double r[2];
    FV_HIP_CHECK(ctx, hipMemcpyAsync(r, cl->verr, sizeof r, hipMemcpyDeviceToHost, ctx->stream));
    FV_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    *err = r[1] > 0 ? r[0] / r[1] : r[0];
    if (!std::isfinite(*err)) *err = INFINITY;
    return FV_OK;
}

int closure_step(fv_vpm *f)
{
    fv_closure *cl = f->closure;
    if (!cl) return FV_OK;
    cl->steps++;   // this advect step; its stages reuse the binning made here
    cl->built_epoch = -1;
    int rc = closure_build(f, false, false);
    if (rc) return rc;
    rc = closure_train(f);
    if (rc) return rc;
    if ((cl->steps - 1) % cl->check_every == 0) {
        double err = 0.0;
        rc = closure_check(f, true, &err);
        if (rc) return rc;
        cl->last_err = err;
        cl->checks++;
        cl->active = err <= cl->tol;
        if (!cl->active) cl->rejections++;
    }
    return FV_OK;
}

bool closure_active(const fv_vpm *f) { return f->closure && f->closure->active; }

int closure_self_sum(fv_vpm *f, int mode, int kernel, double *out)
{
    return closure_pass(f, mode, kernel, out, true, true);
}

void closure_destroy(fv_closure *cl)
{
    if (!cl) return;
    if (cl->side) {
        (void)hipStreamSynchronize(cl->side);
        (void)hipEventDestroy(cl->ev_bin);
        (void)hipEventDestroy(cl->ev_train);
        (void)hipStreamDestroy(cl->side);
    }
    void *bufs[] = {cl->block, cl->cells, cl->far, cl->band, cl->list, cl->tmp, cl->bbox, cl->dstencil,
                    cl->tX, cl->vdir, cl->lpart};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    delete cl;
}

}  // namespace fv

using namespace fv;

extern "C" {

int fv_vpm_set_closure(fv_vpm *f, fv_mlp *mlp, double cell, double kappa, int32_t check_every,
                       double tol, int32_t train_samples, int32_t train_iters, double lr)
{
    if (!f) return FV_E_ARG;
    fv_ctx *ctx = f->ctx;
    FV_REQUIRE(ctx, cell > 0 && kappa >= 2.0 && check_every >= 1 && tol > 0 &&
                        train_samples >= 16 && train_iters >= 0 && lr >= 0,
               "invalid closure parameters");
    closure_destroy(f->closure);
    fv_closure *cl = new (std::nothrow) fv_closure();
    if (!cl) return set_error(ctx, FV_E_HIP, "out of host memory");
    cl->mlp = mlp;
    cl->h = cell;
    cl->kappa = kappa;
    cl->check_every = check_every;
    cl->tol = tol;
    cl->train_samples = (train_samples + 15) / 16 * 16;
    cl->train_iters = train_iters;
    cl->lr = lr;
    f->closure = cl;
    return FV_OK;
}

int fv_vpm_set_closure_local(fv_vpm *f, double kappa_local)
{
    if (!f) return FV_E_ARG;
    fv_ctx *ctx = f->ctx;
    FV_REQUIRE(ctx, f->closure, "no closure enabled");
    fv_closure *cl = f->closure;
    FV_REQUIRE(ctx, kappa_local == 0.0 || kappa_local >= kShell * cl->kappa,
               "kappa_local %g must be 0 or at least %g (outside the learned shell)", kappa_local,
               kShell * cl->kappa);
    cl->kappa_local = kappa_local;
    cl->built_epoch = -1;   // a full rebuild: band lists and local expansions
    cl->binned_step = -1;
    return FV_OK;
}

int fv_vpm_clear_closure(fv_vpm *f)
{
    if (!f) return FV_E_ARG;
    (void)hipStreamSynchronize(f->ctx->stream);
    closure_destroy(f->closure);
    f->closure = nullptr;
    return FV_OK;
}

int fv_vpm_closure_near_pairs(fv_vpm *f, double *pairs)
{
    if (!f || !f->closure || !pairs) return FV_E_ARG;
    fv_ctx *ctx = f->ctx;
    unsigned long long v[2] = {0, 0};
    if (f->closure->near_cnt) {
        FV_HIP_CHECK(ctx, hipMemcpyAsync(v, f->closure->near_cnt, sizeof v, hipMemcpyDeviceToHost,
                                         ctx->stream));
        FV_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    }
    *pairs = (double)v[1];
    return FV_OK;
}

int fv_vpm_closure_stats(fv_vpm *f, double *last_err, int64_t *checks, int64_t *rejections,
                         int32_t *active, double *far_evals)
{
    if (!f || !f->closure) return FV_E_ARG;
    const fv_closure *cl = f->closure;
    if (last_err) *last_err = cl->last_err;
    if (checks) *checks = cl->checks;
    if (rejections) *rejections = cl->rejections;
    if (active) *active = cl->active ? 1 : 0;
    if (far_evals) *far_evals = cl->far_evals;
    return FV_OK;
}

// Closure velocity at the particles (host, original order) with or without
// the learned residual, and its normwise error against the exact sum.
int fv_vpm_closure_velocity_host(fv_vpm *f, int32_t use_mlp, double *out, double *err)
{
    if (!f || !f->closure) return FV_E_ARG;
    fv_ctx *ctx = f->ctx;
    double e = 0.0;
    int rc = closure_check(f, use_mlp != 0, &e);
    if (rc) return rc;
    if (err) *err = e;
    if (out && f->closure->n)
        FV_HIP_CHECK(ctx, hipMemcpyAsync(out, f->closure->vclo,
                                         sizeof(double) * 3 * f->closure->n,
                                         hipMemcpyDeviceToHost, ctx->stream));
    FV_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return FV_OK;
}

// One closure pass (mode 0 velocity, 2 stretching J alpha, 3 vorticity) at the
// particles, host copy in the original order.
int fv_vpm_closure_sum_host(fv_vpm *f, int32_t mode, int32_t kernel, int32_t use_mlp,
                            double *out)
{
    if (!f || !f->closure || !out) return FV_E_ARG;
    fv_ctx *ctx = f->ctx;
    FV_REQUIRE(ctx, mode == 0 || mode == 2 || mode == 3 || mode == 4,
               "closure mode must be 0, 2, 3 or 4");
    FV_REQUIRE(ctx, kernel == FV_KERNEL_GAUSSIAN_ERF || kernel == FV_KERNEL_WINCKELMANS,
               "unknown kernel");
    fv_closure *cl = f->closure;
    if (!cl->vdir) {
        FV_HIP_CHECK(ctx, hipMalloc(&cl->vdir, sizeof(double) * (6 * (size_t)f->cap + 8)));
        cl->vclo = cl->vdir + 3 * (size_t)f->cap;
        cl->verr = cl->vclo + 3 * (size_t)f->cap;
    }
    int rc = closure_pass(f, mode, kernel, cl->vclo, use_mlp != 0);
    if (rc) return rc;
    if (cl->n)
        FV_HIP_CHECK(ctx, hipMemcpyAsync(out, cl->vclo, sizeof(double) * 3 * cl->n,
                                         hipMemcpyDeviceToHost, ctx->stream));
    FV_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return FV_OK;
}

// Runs the closure's online training once on the current particle state.
int fv_vpm_closure_train(fv_vpm *f, int32_t rounds)
{
    if (!f || !f->closure) return FV_E_ARG;
    int rc = closure_build(f, false, false);
    for (int r = 0; !rc && r < rounds; ++r) rc = closure_train(f);
    if (!rc) rc = join_training(f->closure, f->ctx);
    if (!rc) rc = fv_ctx_synchronize(f->ctx);
    return rc;
}

}  // extern "C"
