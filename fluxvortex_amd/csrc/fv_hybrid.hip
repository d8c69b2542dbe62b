// fv_hybrid.hip -- device-resident hybrid panel-particle UVLM step.
//
// Reference: tests/experiment_hybrid_panel_particle.py:87-204 (HybridSolver)
// over PteraSoftware's UnsteadyRingVortexLatticeMethodSolver step loop
// (idea-stage/architecture_analysis.md:5-23).  Per step, all on the context
// stream with no host synchronisation:
//   geometry   ring vortices at quarter chord, collocation points at 3/4
//              chord, unit normals, areas, leg centres, motion velocities
//   AIC        unit-strength rings . normals                       (P x P)
//   rhs        -[(V - v_motion).n + (v_wake_rings + v_particles).n]
//   solve      dense LU (FP64) or FP32 LU + FP64 refinement (mixed)
//   loads      Kutta-Joukowski on right/front/left legs (and the trailing-edge
//              back leg against the newest wake ring) with effective
//              strengths + unsteady -rho A dGamma/dt n  ->  CL, CDi
//   wake       shed the trailing-edge row, advect the ring wake with V dt
//   convert    ring row n_keep -> 4 particles per ring, ring gamma := 0
//   particles  RK3 advection (+ stretching) + Pedrizzetti relaxation
// Wake rows beyond n_keep hold zero circulation after conversion and are not
// stored (they add exact zeros to every sum).
#include <cmath>
#include <cstring>
#include <new>
#include <vector>

#include "fv_lu.h"
#include "fv_vpm.h"

namespace fv {
int dense_solve(fv_ctx *ctx, int64_t batch, int64_t n, double *A, double *rhs, int32_t *info,
                const LuReuse *reuse, bool store_factors);
int dense_solve_mixed(fv_ctx *ctx, int64_t batch, int64_t n, const double *A, double *rhs,
                      int max_iter, double tol, int32_t *info, int32_t *iters,
                      const LuReuse *reuse);
}

struct fv_surface {
    int nc, ns;
    int64_t panel_off;   // first panel of this surface
    int64_t vert_off;    // first grid vertex of this surface (within one step)
    int64_t wvert_off;   // first wake vertex (within one wake buffer)
    int64_t wring_off;   // first wake ring slot (within one wake buffer)
};

struct fv_hybrid {
    fv_ctx *ctx = nullptr;
    fv_hybrid_desc d{};
    std::vector<fv_surface> surf;
    int64_t P = 0, TV = 0;       // panels, grid vertices per step
    int64_t R = 0;               // stored wake ring rows per surface
    int64_t WV = 0, WR = 0;      // wake vertex / ring slots (all surfaces)
    int64_t NS = 0;              // total spanwise panels (rings per wake row)
    int64_t step = 0;            // next step to run
    int64_t nrow = 0;            // ring rows currently in the wake
    int par = 0;                 // wake double-buffer parity
    hipStream_t side = nullptr;  // the circulation solve's stream (overlaps the advection)
    hipEvent_t ev_rhs = nullptr, ev_solved = nullptr;
    void *block = nullptr;
    double *grids = nullptr;
    double *br = nullptr, *fr = nullptr, *fl = nullptr, *bl = nullptr;
    double *colloc = nullptr, *normal = nullptr, *area = nullptr, *vm_col = nullptr;
    double *leg_c = nullptr, *leg_v = nullptr, *vm_leg = nullptr;
    double *aic = nullptr, *rhs = nullptr, *gam = nullptr, *gam_prev = nullptr;
    double *vw = nullptr, *vp = nullptr, *vl = nullptr;
    double *wv[2] = {nullptr, nullptr}, *wg[2] = {nullptr, nullptr}, *wa[2] = {nullptr, nullptr};
    double *rbr = nullptr, *rfr = nullptr, *rfl = nullptr, *rbl = nullptr;   // wake ring corners
    double *rg = nullptr, *rage = nullptr, *rrc0 = nullptr, *brc0 = nullptr;
    double *cbr = nullptr, *cfr = nullptr, *cfl = nullptr, *cbl = nullptr, *cg = nullptr;
    double *cl = nullptr, *cdi = nullptr, *force = nullptr;
    fv::SegRec *unit_tab = nullptr, *tab = nullptr;
    int32_t *info = nullptr;                           // dense-solve status
    int64_t *step_dev = nullptr;                       // the step counter kernels read
    // factor reuse of the circulation solve: a step whose AIC equals the last
    // factored one bit for bit (a lattice that does not move) skips the
    // factorization (the solution is the same bit for bit)
    fv::LuReuse reuse{nullptr, nullptr, nullptr, nullptr, nullptr};
    int reuse_on = 1;
    int32_t *d_nc = nullptr, *d_ns = nullptr;          // per-surface sizes (device)
    int64_t *d_poff = nullptr;                         // per-surface panel offsets (device)
    int64_t *d_woff = nullptr;                         // per-surface wake ring offsets (device)
    fv_vpm *vpm = nullptr;
    // population control after each advect (0 none, 1 drop oldest, 2 merge)
    int pop_scheme = 0;
    // free ring wake: vertices move with V + the bound / wake-ring induction
    // (frozen snapshot, forward Euler); rings beyond trunc_x get gamma = 0
    int ring_free = 0;
    double trunc_x = INFINITY;
    double *wpts = nullptr, *wvel = nullptr;   // 3 WV each
    int64_t pop_max = 0, pop_protect = 64;
    double pop_eps = 1e-3, pop_protect_dist = 0.0;
    double seg_pairs = 0.0;
    // graph mode: launches sized by capacities and device counts, so every
    // steady step (below) enqueues the same work and replays from a graph
    // (1), or the same launches run eagerly (2)
    int graph_mode = 0;
    bool warm = false;               // a steady step ran eagerly in graph mode
    fv_hybrid_batch *self = nullptr; // the one-run batch fv_hybrid_step replays through
};

// Host-side state a step advances (restored after a capture, which runs the
// host code of one step without executing it).
struct HybridHost {
    int64_t step, nrow;
    int par;
    double seg_pairs;
    int64_t np_upper, np_host, state_epoch;
    bool np_exact;
};

// Everything the enqueued work of a steady step depends on besides the
// device state; a change makes the captured graphs stale.
struct HybridSig {
    int32_t free_wake, stretch, precision, pop_scheme, ring_free, reuse_on;
    int32_t kernel, vprec, pair_rule;
    double trunc_x, rlxf, nu;
    const void *closure;
    bool operator==(const HybridSig &o) const { return memcmp(this, &o, sizeof o) == 0; }
};

struct fv_hybrid_batch {
    std::vector<fv_hybrid *> runs;
    hipStream_t master = nullptr;
    hipEvent_t ev_fork = nullptr;
    std::vector<hipEvent_t> ev_join;
    hipGraphExec_t exec[2] = {nullptr, nullptr};   // by wake parity
    std::vector<HybridHost> before[2], after[2];   // each run's host state around the capture
    std::vector<HybridSig> sig[2];
    bool on_master = false;   // the runs' streams wait for the master's graph launches
    int64_t graph_launches = 0, eager_steps = 0;
};

namespace fv {

__device__ __forceinline__ void lerp3(const double *a, const double *b, double t, double *o)
{
    o[0] = a[0] + t * (b[0] - a[0]);
    o[1] = a[1] + t * (b[1] - a[1]);
    o[2] = a[2] + t * (b[2] - a[2]);
}

// Ring vertices of panel (i, j) from a (nc+1) x (ns+1) corner grid g:
// front legs at the panel's quarter chord; back legs at the next panel's
// quarter chord, or one panel length behind the front leg on the last row.
__device__ void panel_ring(const double *g, int nc, int ns, int i, int j, double *Br, double *Fr,
                           double *Fl, double *Bl)
{
    const double *v00 = g + 3 * ((int64_t)i * (ns + 1) + j);
    const double *v01 = g + 3 * ((int64_t)i * (ns + 1) + j + 1);
    const double *v10 = g + 3 * ((int64_t)(i + 1) * (ns + 1) + j);
    const double *v11 = g + 3 * ((int64_t)(i + 1) * (ns + 1) + j + 1);
    lerp3(v00, v10, 0.25, Fl);
    lerp3(v01, v11, 0.25, Fr);
    if (i < nc - 1) {
        const double *w0 = g + 3 * ((int64_t)(i + 2) * (ns + 1) + j);
        const double *w1 = g + 3 * ((int64_t)(i + 2) * (ns + 1) + j + 1);
        lerp3(v10, w0, 0.25, Bl);
        lerp3(v11, w1, 0.25, Br);
    } else {
        for (int c = 0; c < 3; ++c) {
            Bl[c] = Fl[c] + (v10[c] - v00[c]);
            Br[c] = Fr[c] + (v11[c] - v01[c]);
        }
    }
}

// Trailing-edge ring back vertex at span station j (the wake's front row).
__device__ void te_vertex(const double *g, int nc, int ns, int j, double *o)
{
    const double *a = g + 3 * ((int64_t)(nc - 1) * (ns + 1) + j);
    const double *b = g + 3 * ((int64_t)nc * (ns + 1) + j);
    for (int c = 0; c < 3; ++c) o[c] = b[c] + 0.25 * (b[c] - a[c]);
}

__device__ void panel_frame(const double *g, int ns, int i, int j, double *col, double *nrm,
                            double *area)
{
    const double *v00 = g + 3 * ((int64_t)i * (ns + 1) + j);        // front-left
    const double *v01 = g + 3 * ((int64_t)i * (ns + 1) + j + 1);    // front-right
    const double *v10 = g + 3 * ((int64_t)(i + 1) * (ns + 1) + j);  // back-left
    const double *v11 = g + 3 * ((int64_t)(i + 1) * (ns + 1) + j + 1);
    double a[3], b[3];
    lerp3(v00, v10, 0.75, a);
    lerp3(v01, v11, 0.75, b);
    for (int c = 0; c < 3; ++c) col[c] = 0.5 * (a[c] + b[c]);
    // diagonals (Fr - Bl) x (Fl - Br): +z for a wing with x aft, y right
    const double d1[3] = {v01[0] - v10[0], v01[1] - v10[1], v01[2] - v10[2]};
    const double d2[3] = {v00[0] - v11[0], v00[1] - v11[1], v00[2] - v11[2]};
    const double cx = d1[1] * d2[2] - d1[2] * d2[1];
    const double cy = d1[2] * d2[0] - d1[0] * d2[2];
    const double cz = d1[0] * d2[1] - d1[1] * d2[0];
    const double len = sqrt(cx * cx + cy * cy + cz * cz);
    *area = 0.5 * len;
    nrm[0] = cx / len;
    nrm[1] = cy / len;
    nrm[2] = cz / len;
}

// Leg l (0 right Br->Fr, 1 front Fr->Fl, 2 left Fl->Bl, 3 back Bl->Br): centre
// and vector.
__device__ void leg_of(const double *Br, const double *Fr, const double *Fl, const double *Bl,
                       int l, double *c, double *v)
{
    const double *s = l == 0 ? Br : (l == 1 ? Fr : (l == 2 ? Fl : Bl));
    const double *e = l == 0 ? Fr : (l == 1 ? Fl : (l == 2 ? Bl : Br));
    for (int k = 0; k < 3; ++k) {
        c[k] = 0.5 * (s[k] + e[k]);
        v[k] = e[k] - s[k];
    }
}

// The corner grids of every step (step-major, `stride` doubles a step) and the
// device step counter: kernels find this step's grid (and its neighbours)
// themselves, so a step's launches do not change from step to step.
struct StepGrid {
    const double *g;
    int64_t stride;
    const int64_t *step;
    // grid of step (current + delta), offset by off doubles; null before step 0
    __device__ const double *at(int delta, int64_t off) const
    {
        const int64_t s = *step + delta;
        return s < 0 ? nullptr : g + s * stride + off;
    }
};

__global__ void k_geometry(StepGrid sg, int64_t voff, int nc, int ns, int64_t poff, double inv_dt,
                           double *__restrict__ br, double *__restrict__ fr,
                           double *__restrict__ fl, double *__restrict__ bl,
                           double *__restrict__ colloc, double *__restrict__ normal,
                           double *__restrict__ area, double *__restrict__ vm_col,
                           double *__restrict__ leg_c, double *__restrict__ leg_v,
                           double *__restrict__ vm_leg)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nc * ns) return;
    const double *gcur = sg.at(0, 3 * voff), *gprev = sg.at(-1, 3 * voff);
    const int i = t / ns, j = t % ns;
    const int64_t p = poff + t;
    double Br[3], Fr[3], Fl[3], Bl[3], col[3], nrm[3], ar;
    panel_ring(gcur, nc, ns, i, j, Br, Fr, Fl, Bl);
    panel_frame(gcur, ns, i, j, col, nrm, &ar);
    for (int c = 0; c < 3; ++c) {
        br[3 * p + c] = Br[c];
        fr[3 * p + c] = Fr[c];
        fl[3 * p + c] = Fl[c];
        bl[3 * p + c] = Bl[c];
        colloc[3 * p + c] = col[c];
        normal[3 * p + c] = nrm[c];
    }
    area[p] = ar;
    double pBr[3], pFr[3], pFl[3], pBl[3], pcol[3], pn[3], pa;
    if (gprev) {
        panel_ring(gprev, nc, ns, i, j, pBr, pFr, pFl, pBl);
        panel_frame(gprev, ns, i, j, pcol, pn, &pa);
    }
    for (int c = 0; c < 3; ++c) vm_col[3 * p + c] = gprev ? (col[c] - pcol[c]) * inv_dt : 0.0;
    for (int l = 0; l < 4; ++l) {
        double lc[3], lv[3];
        leg_of(Br, Fr, Fl, Bl, l, lc, lv);
        double plc[3], plv[3];
        if (gprev) leg_of(pBr, pFr, pFl, pBl, l, plc, plv);
        for (int c = 0; c < 3; ++c) {
            leg_c[3 * (4 * p + l) + c] = lc[c];
            leg_v[3 * (4 * p + l) + c] = lv[c];
            vm_leg[3 * (4 * p + l) + c] = gprev ? (lc[c] - plc[c]) * inv_dt : 0.0;
        }
    }
}

// rhs_i = -[ (dot(v_w, n) + dot(v_p, n)) + dot(V - v_m, n) ]
__global__ void k_rhs(int64_t P, const double *__restrict__ normal,
                      const double *__restrict__ vm_col, const double *__restrict__ vw,
                      const double *__restrict__ vp, double vx, double vy, double vz,
                      double *__restrict__ rhs)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P) return;
    const double *n = normal + 3 * i;
    const double fs = (vx - vm_col[3 * i]) * n[0] + (vy - vm_col[3 * i + 1]) * n[1] +
                      (vz - vm_col[3 * i + 2]) * n[2];
    const double w = vw[3 * i] * n[0] + vw[3 * i + 1] * n[1] + vw[3 * i + 2] * n[2];
    const double q = vp ? (vp[3 * i] * n[0] + vp[3 * i + 1] * n[1] + vp[3 * i + 2] * n[2]) : 0.0;
    rhs[i] = -((w + q) + fs);
}

// One workgroup: per-panel forces in panel order, fixed-order reduction.
// Kutta-Joukowski on each leg with the net circulation of its line: right /
// left legs against the spanwise neighbour, the front leg against the ring
// ahead (whole on the leading edge), the back leg only on the trailing edge,
// against the newest wake ring (whole when there is none); interior back legs
// are the next ring's front legs.
constexpr int kLoadBlock = 256;

__global__ __launch_bounds__(kLoadBlock) void k_loads(
    int nsurf, const int *__restrict__ surf_nc, const int *__restrict__ surf_ns,
    const int64_t *__restrict__ surf_poff, const int64_t *__restrict__ surf_woff,
    const double *__restrict__ wake_g, int64_t nrow, int64_t P, const double *__restrict__ gam,
    const double *__restrict__ gam_prev, const double *__restrict__ leg_v,
    const double *__restrict__ vm_leg, const double *__restrict__ vl,
    const double *__restrict__ normal, const double *__restrict__ area, double vx, double vy,
    double vz, double rho, double inv_dt, double sref, const int64_t *__restrict__ step_dev,
    double *__restrict__ panel_force, double *__restrict__ cl, double *__restrict__ cdi)
{
    __shared__ double red[kLoadBlock / 64][3];
    const int64_t step = *step_dev;
    const bool first_step = step == 0;
    double acc[3] = {0.0, 0.0, 0.0};
    for (int64_t p = threadIdx.x; p < P; p += kLoadBlock) {
        int k = 0;
        while (k + 1 < nsurf && p >= surf_poff[k + 1]) ++k;
        const int ns = surf_ns[k], nc = surf_nc[k];
        const int64_t loc = p - surf_poff[k];
        const int i = (int)(loc / ns), j = (int)(loc % ns);
        const double G = gam[p];
        double geff[4];
        geff[0] = (j + 1 < ns) ? G - gam[p + 1] : G;    // right leg
        geff[1] = (i > 0) ? G - gam[p - ns] : G;         // front leg
        geff[2] = (j > 0) ? G - gam[p - 1] : G;          // left leg
        geff[3] = G - (nrow > 0 ? wake_g[surf_woff[k] + j] : 0.0);   // trailing-edge back leg
        const int nleg = i == nc - 1 ? 4 : 3;
        double F[3] = {0.0, 0.0, 0.0};
        for (int l = 0; l < nleg; ++l) {
            const int64_t q = 4 * p + l;
            const double V[3] = {vx - vm_leg[3 * q] + vl[3 * q], vy - vm_leg[3 * q + 1] + vl[3 * q + 1],
                                 vz - vm_leg[3 * q + 2] + vl[3 * q + 2]};
            const double *L = leg_v + 3 * q;
            const double f = rho * geff[l];
            F[0] += f * (V[1] * L[2] - V[2] * L[1]);
            F[1] += f * (V[2] * L[0] - V[0] * L[2]);
            F[2] += f * (V[0] * L[1] - V[1] * L[0]);
        }
        const double dG = first_step ? 0.0 : (G - gam_prev[p]) * inv_dt;
        const double fu = -rho * area[p] * dG;
        for (int c = 0; c < 3; ++c) {
            F[c] += fu * normal[3 * p + c];
            panel_force[3 * p + c] = F[c];
            acc[c] += F[c];
        }
    }
    for (int c = 0; c < 3; ++c)
        for (int off = 32; off > 0; off >>= 1) acc[c] += __shfl_xor(acc[c], off, 64);
    if ((threadIdx.x & 63) == 0)
        for (int c = 0; c < 3; ++c) red[threadIdx.x / 64][c] = acc[c];
    __syncthreads();
    if (threadIdx.x == 0) {
        double F[3] = {0.0, 0.0, 0.0};
        for (int w = 0; w < kLoadBlock / 64; ++w)
            for (int c = 0; c < 3; ++c) F[c] += red[w][c];
        const double V2 = vx * vx + vy * vy + vz * vz;
        const double Vn = sqrt(V2);
        const double d[3] = {vx / Vn, vy / Vn, vz / Vn};
        // lift direction d x y_hat, normalised
        double l[3] = {-d[2], 0.0, d[0]};
        const double ln = sqrt(l[0] * l[0] + l[2] * l[2]);
        l[0] /= ln;
        l[2] /= ln;
        const double qs = 0.5 * rho * V2 * sref;
        cl[step] = (F[0] * l[0] + F[1] * l[1] + F[2] * l[2]) / qs;
        cdi[step] = (F[0] * d[0] + F[1] * d[1] + F[2] * d[2]) / qs;
    }
}

__global__ void k_copy(const double *__restrict__ a, double *__restrict__ b, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = a[i];
}

// Old wake vertex rows 0 .. nr-1 of one surface as points: row 0 = trailing
// edge of the current geometry, row r = old row r.
__global__ void k_wake_points(StepGrid sg, int64_t voff, int nc, int ns,
                              const double *__restrict__ wold, int64_t nr,
                              double *__restrict__ pts)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nr * (ns + 1)) return;
    const double *gcur = sg.at(0, 3 * voff);
    const int64_t r = t / (ns + 1);
    const int j = (int)(t % (ns + 1));
    double v[3];
    if (r == 0) te_vertex(gcur, nc, ns, j, v);
    else
        for (int c = 0; c < 3; ++c) v[c] = wold[3 * t + c];
    for (int c = 0; c < 3; ++c) pts[3 * t + c] = v[c];
}

// New wake vertex rows of one surface: row 0 = trailing edge of the next
// geometry, row r+1 = (r == 0 ? trailing edge of the current geometry : old
// row r) + (V + u_r) dt for r < nrow_new, with u_r the induced velocity at
// that old vertex (vel; nullptr for the prescribed wake, u = 0).
__global__ void k_wake_vertices(StepGrid sg, int64_t voff, int nc, int ns,
                                const double *__restrict__ wold,
                                const double *__restrict__ vel, double *__restrict__ wnew,
                                int64_t nrow_new, double vx, double vy, double vz, double dt)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (nrow_new + 1) * (ns + 1)) return;
    const double *gnext = sg.at(1, 3 * voff), *gcur = sg.at(0, 3 * voff);
    const int64_t r = t / (ns + 1);
    const int j = (int)(t % (ns + 1));
    double v[3];
    if (r == 0) {
        te_vertex(gnext, nc, ns, j, v);
    } else {
        if (r == 1) te_vertex(gcur, nc, ns, j, v);
        else
            for (int c = 0; c < 3; ++c) v[c] = wold[3 * ((r - 1) * (ns + 1) + j) + c];
        const double *u = vel ? vel + 3 * ((r - 1) * (ns + 1) + j) : nullptr;
        v[0] += (u ? vx + u[0] : vx) * dt;
        v[1] += (u ? vy + u[1] : vy) * dt;
        v[2] += (u ? vz + u[2] : vz) * dt;
    }
    for (int c = 0; c < 3; ++c) wnew[3 * t + c] = v[c];
}

// Rings in rows [r0, r1) of one surface's new wake whose corner-mean x
// exceeds max_x lose their circulation (warp_fsi/kernels_wake.py:110-126).
__global__ void k_wake_truncate(const double *__restrict__ w, int ns, int64_t r0, int64_t r1,
                                double max_x, double *__restrict__ g)
{
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (r1 - r0) * ns) return;
    const int64_t r = r0 + q / ns;
    const int j = (int)(q % ns);
    const int64_t t = r * ns + j;
    const double *a = w + 3 * (r * (ns + 1) + j), *b = a + 3;
    const double *c = w + 3 * ((r + 1) * (ns + 1) + j), *e = c + 3;
    const double cx = (b[0] + a[0] + c[0] + e[0]) * 0.25;
    if (cx > max_x) g[t] = 0.0;
}

// New ring strengths / ages of rows [r0, r1) of one surface: row 0 =
// trailing-edge bound circulations (age 0), row r+1 = old row r (age + dt).
__global__ void k_wake_rings(const double *__restrict__ gam_te, int ns,
                             const double *__restrict__ gold, const double *__restrict__ aold,
                             double *__restrict__ gnew, double *__restrict__ anew, int64_t r0,
                             int64_t r1, double dt)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (r1 - r0) * ns) return;
    const int64_t r = r0 + e / ns;
    const int j = (int)(e % ns);
    const int64_t t = r * ns + j;
    if (r == 0) {
        gnew[t] = gam_te[j];
        anew[t] = 0.0;
    } else {
        gnew[t] = gold[(r - 1) * ns + j];
        anew[t] = aold[(r - 1) * ns + j] + dt;
    }
}

// Ring corners of rows [r0, r0 + nr) of one surface's wake, written at out + off.
__global__ void k_wake_corners(const double *__restrict__ w, const double *__restrict__ g,
                               const double *__restrict__ a, int ns, int64_t r0, int64_t nr,
                               int64_t off, double *__restrict__ br, double *__restrict__ fr,
                               double *__restrict__ fl, double *__restrict__ bl,
                               double *__restrict__ og, double *__restrict__ oa)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nr * ns) return;
    const int64_t r = r0 + t / ns;
    const int j = (int)(t % ns);
    const int64_t o = off + t;
    const double *vfl = w + 3 * (r * (ns + 1) + j);
    const double *vfr = w + 3 * (r * (ns + 1) + j + 1);
    const double *vbl = w + 3 * ((r + 1) * (ns + 1) + j);
    const double *vbr = w + 3 * ((r + 1) * (ns + 1) + j + 1);
    for (int c = 0; c < 3; ++c) {
        fl[3 * o + c] = vfl[c];
        fr[3 * o + c] = vfr[c];
        bl[3 * o + c] = vbl[c];
        br[3 * o + c] = vbr[c];
    }
    if (og) og[o] = g[r * ns + j];
    if (oa) oa[o] = a ? a[r * ns + j] : 0.0;
}

// The device step counter moves to the next step (the step's last launch).
__global__ void k_step_advance(int64_t *__restrict__ step) { *step += 1; }

// Writes converted strengths back into one surface's wake row.
__global__ void k_scatter_row(const double *__restrict__ src, int ns, int64_t row,
                              double *__restrict__ g)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < ns) g[row * ns + j] = src[j];
}

static int hybrid_step_one(fv_hybrid *h)
{
    fv_ctx *ctx = h->ctx;
    const fv_hybrid_desc &d = h->d;
    const int64_t s = h->step;
    const int64_t P = h->P;
    const double dt = d.dt;
    const StepGrid sg{h->grids, h->TV * 3, h->step_dev};
    int rc;
    // geometry
    for (auto &sf : h->surf) {
        const int np = sf.nc * sf.ns;
        hipLaunchKernelGGL(k_geometry, dim3((unsigned)ceil_div(np, 128)), dim3(128), 0,
                           ctx->stream, sg, sf.vert_off, sf.nc, sf.ns, sf.panel_off,
                           1.0 / dt, h->br, h->fr, h->fl, h->bl, h->colloc, h->normal, h->area,
                           h->vm_col, h->leg_c, h->leg_v, h->vm_leg);
        FV_LAUNCH_CHECK(ctx);
    }
    // AIC from unit-strength bound rings
    rc = build_segtab(ctx, h->br, h->fr, h->fl, h->bl, P, nullptr, h->brc0, nullptr, d.nu,
                      FV_LEGS_RING, true, h->unit_tab);
    if (rc) return rc;
    rc = aic_assemble(ctx, h->colloc, h->normal, P, h->unit_tab, h->aic);
    if (rc) return rc;
    h->seg_pairs += 4.0 * P * P;
    // wake ring corners for all stored rows
    const int cur = h->par;
    int64_t W = 0;
    for (auto &sf : h->surf) {
        if (h->nrow > 0) {
            hipLaunchKernelGGL(k_wake_corners, dim3((unsigned)ceil_div(h->nrow * sf.ns, 128)),
                               dim3(128), 0, ctx->stream, h->wv[cur] + 3 * sf.wvert_off,
                               h->wg[cur] + sf.wring_off, h->wa[cur] + sf.wring_off, sf.ns,
                               (int64_t)0, h->nrow, W, h->rbr, h->rfr, h->rfl, h->rbl, h->rg,
                               h->rage);
            FV_LAUNCH_CHECK(ctx);
        }
        W += h->nrow * sf.ns;
    }
    // segment table: bound rings (filled after the solve) | wake rings
    SegRec *wake_tab = h->tab + 4 * P;
    if (W > 0) {
        rc = build_segtab(ctx, h->rbr, h->rfr, h->rfl, h->rbl, W, h->rg, h->rrc0, h->rage, d.nu,
                          FV_LEGS_RING, false, wake_tab);
        if (rc) return rc;
    }
    const int prec = d.precision;
    rc = seg_sum(ctx, h->colloc, P, wake_tab, 4 * W, prec, h->vw, S_PARTIAL);
    if (rc) return rc;
    h->seg_pairs += 4.0 * P * W;
    const double *vp = nullptr;
    if (h->vpm->np_upper > 0) {
        rc = vpm_induce(h->vpm, h->colloc, P, prec, h->vp, S_PARTIAL);
        if (rc) return rc;
        vp = h->vp;
    }
    hipLaunchKernelGGL(k_rhs, dim3((unsigned)ceil_div(P, 128)), dim3(128), 0, ctx->stream, P,
                       h->normal, h->vm_col, h->vw, vp, d.vinf[0], d.vinf[1], d.vinf[2], h->rhs);
    FV_LAUNCH_CHECK(ctx);
    // circulation solve, loads, shedding, conversion, advection.  When the ring
    // row that converts this step is at least one step old (n_keep >= 1) and the
    // ring wake is prescribed, only the loads and the newest wake row need this
    // step's circulation: the solve then runs on the side stream while this
    // stream sheds the older rows, converts and advects the particles, and the
    // loads and the newest row follow once the solve is done (the same kernels
    // on the same inputs as the sequential order, so the results are identical).
    const bool shed = s + 1 < d.num_steps;
    const bool overlap = shed && d.n_keep >= 1 && !h->ring_free && h->side;
    const int64_t nrow_old = h->nrow;
    auto solve = [&]() -> int {
        const LuReuse *ru = h->reuse_on ? &h->reuse : nullptr;
        if (prec == FV_PREC_FP64) return dense_solve(ctx, 1, P, h->aic, h->rhs, h->info, ru, false);
        return dense_solve_mixed(ctx, 1, P, h->aic, h->rhs, 8, 1e-12, h->info, nullptr, ru);
    };
    // loads: velocities at leg centres from bound + wake rings (this step's wake)
    auto loads = [&]() -> int {
        hipLaunchKernelGGL(k_copy, dim3((unsigned)ceil_div(P, 128)), dim3(128), 0, ctx->stream,
                           h->rhs, h->gam, P);
        FV_LAUNCH_CHECK(ctx);
        int r = build_segtab(ctx, h->br, h->fr, h->fl, h->bl, P, h->gam, h->brc0, nullptr, d.nu,
                             FV_LEGS_RING, false, h->tab);
        if (r) return r;
        r = seg_sum(ctx, h->leg_c, 4 * P, h->tab, 4 * P + 4 * W, prec, h->vl, S_PARTIAL);
        if (r) return r;
        h->seg_pairs += 4.0 * P * (4.0 * P + 4.0 * W);
        hipLaunchKernelGGL(k_loads, dim3(1), dim3(kLoadBlock), 0, ctx->stream,
                           (int)h->surf.size(), h->d_nc, h->d_ns, h->d_poff, h->d_woff, h->wg[cur],
                           nrow_old, P, h->gam, h->gam_prev, h->leg_v, h->vm_leg,
                           h->vl, h->normal, h->area, d.vinf[0], d.vinf[1], d.vinf[2], d.rho,
                           1.0 / dt, d.sref, h->step_dev, h->force, h->cl, h->cdi);
        FV_LAUNCH_CHECK(ctx);
        hipLaunchKernelGGL(k_copy, dim3((unsigned)ceil_div(P, 128)), dim3(128), 0, ctx->stream,
                           h->gam, h->gam_prev, P);
        FV_LAUNCH_CHECK(ctx);
        return FV_OK;
    };
    // new ring strengths / ages of rows [r0, r1) of the next wake, then the
    // truncation of those rows
    auto wake_rows = [&](int nxt, int64_t r0, int64_t r1) -> int {
        for (auto &sf : h->surf) {
            const double *gam_te = h->gam + sf.panel_off + (int64_t)(sf.nc - 1) * sf.ns;
            hipLaunchKernelGGL(k_wake_rings, dim3((unsigned)ceil_div((r1 - r0) * sf.ns, 128)),
                               dim3(128), 0, ctx->stream, gam_te, sf.ns, h->wg[cur] + sf.wring_off,
                               h->wa[cur] + sf.wring_off, h->wg[nxt] + sf.wring_off,
                               h->wa[nxt] + sf.wring_off, r0, r1, dt);
            FV_LAUNCH_CHECK(ctx);
        }
        if (std::isfinite(h->trunc_x))
            for (auto &sf : h->surf) {
                hipLaunchKernelGGL(k_wake_truncate,
                                   dim3((unsigned)ceil_div((r1 - r0) * sf.ns, 128)), dim3(128), 0,
                                   ctx->stream, h->wv[nxt] + 3 * sf.wvert_off, sf.ns, r0, r1,
                                   h->trunc_x, h->wg[nxt] + sf.wring_off);
                FV_LAUNCH_CHECK(ctx);
            }
        return FV_OK;
    };
    if (overlap) {
        FV_HIP_CHECK(ctx, hipEventRecord(h->ev_rhs, ctx->stream));
        FV_HIP_CHECK(ctx, hipStreamWaitEvent(h->side, h->ev_rhs, 0));
        hipStream_t own = ctx->stream;
        ctx->stream = h->side;
        rc = solve();
        ctx->stream = own;
        if (rc) return rc;
        FV_HIP_CHECK(ctx, hipEventRecord(h->ev_solved, h->side));
    } else {
        rc = solve();
        if (rc) return rc;
        rc = loads();
        if (rc) return rc;
    }
    // wake shedding + conversion (not after the last step)
    int nxt = cur;
    if (shed) {
        nxt = 1 - cur;
        const int64_t nrow_new = (h->nrow + 1 < h->R) ? h->nrow + 1 : h->R;
        if (h->ring_free) {
            // frozen snapshot: induced velocity at every old vertex that moves,
            // from the bound rings (this step's gamma) and the wake rings
            for (auto &sf : h->surf) {
                const int64_t np_ = nrow_new * (sf.ns + 1);
                hipLaunchKernelGGL(k_wake_points, dim3((unsigned)ceil_div(np_, 128)), dim3(128), 0,
                                   ctx->stream, sg, sf.vert_off, sf.nc, sf.ns,
                                   h->wv[cur] + 3 * sf.wvert_off, nrow_new,
                                   h->wpts + 3 * sf.wvert_off);
                FV_LAUNCH_CHECK(ctx);
                rc = seg_sum(ctx, h->wpts + 3 * sf.wvert_off, np_, h->tab, 4 * P + 4 * W, prec,
                             h->wvel + 3 * sf.wvert_off, S_PARTIAL);
                if (rc) return rc;
                h->seg_pairs += (double)np_ * (4.0 * P + 4.0 * W);
            }
        }
        for (auto &sf : h->surf) {
            hipLaunchKernelGGL(k_wake_vertices,
                               dim3((unsigned)ceil_div((nrow_new + 1) * (sf.ns + 1), 128)),
                               dim3(128), 0, ctx->stream, sg, sf.vert_off, sf.nc, sf.ns,
                               h->wv[cur] + 3 * sf.wvert_off,
                               h->ring_free ? h->wvel + 3 * sf.wvert_off : nullptr,
                               h->wv[nxt] + 3 * sf.wvert_off, nrow_new, d.vinf[0], d.vinf[1],
                               d.vinf[2], dt);
            FV_LAUNCH_CHECK(ctx);
        }
        // rows 1.. hold older circulations; row 0 (this step's) waits for the
        // solve when it runs on the side stream
        rc = overlap ? (nrow_new > 1 ? wake_rows(nxt, 1, nrow_new) : FV_OK)
                     : wake_rows(nxt, 0, nrow_new);
        if (rc) return rc;
        h->par = nxt;
        h->nrow = nrow_new;
        const int64_t kr = d.n_keep;
        if (kr < nrow_new) {
            // ring row n_keep of every surface, in surface order
            int64_t off = 0;
            for (auto &sf : h->surf) {
                hipLaunchKernelGGL(k_wake_corners, dim3((unsigned)ceil_div(sf.ns, 128)),
                                   dim3(128), 0, ctx->stream, h->wv[nxt] + 3 * sf.wvert_off,
                                   h->wg[nxt] + sf.wring_off, (const double *)nullptr, sf.ns, kr,
                                   (int64_t)1, off, h->cbr, h->cfr, h->cfl, h->cbl, h->cg,
                                   (double *)nullptr);
                FV_LAUNCH_CHECK(ctx);
                off += sf.ns;
            }
            rc = vpm_convert_rings(h->vpm, off, h->cbr, h->cfr, h->cfl, h->cbl, h->cg, d.sigma);
            if (rc) return rc;
            off = 0;
            for (auto &sf : h->surf) {
                hipLaunchKernelGGL(k_scatter_row, dim3((unsigned)ceil_div(sf.ns, 128)), dim3(128),
                                   0, ctx->stream, h->cg + off, sf.ns, kr,
                                   h->wg[nxt] + sf.wring_off);
                FV_LAUNCH_CHECK(ctx);
                off += sf.ns;
            }
        }
    }
    // particle advection + relaxation
    if (h->vpm->np_upper > 0) {
        rc = vpm_advect(h->vpm, dt, d.vinf, d.stretch, d.free_wake);
        if (rc) return rc;
    }
    // population control (newton_pc/adapters/flap.py:466-475)
    if (h->pop_scheme == 1 && h->vpm->np_upper > h->pop_max) {
        rc = vpm_drop_oldest(h->vpm, h->pop_max);
        if (rc) return rc;
    } else if (h->pop_scheme == 2 && h->vpm->np_upper >= 2) {
        const double vn = sqrt(d.vinf[0] * d.vinf[0] + d.vinf[1] * d.vinf[1] + d.vinf[2] * d.vinf[2]);
        rc = vpm_merge_colloc(h->vpm, h->colloc, P, h->pop_eps, vn, h->pop_protect,
                              h->pop_protect_dist);
        if (rc) return rc;
    }
    if (overlap) {
        FV_HIP_CHECK(ctx, hipStreamWaitEvent(ctx->stream, h->ev_solved, 0));
        rc = loads();
        if (rc) return rc;
        rc = wake_rows(nxt, 0, 1);
        if (rc) return rc;
    }
    hipLaunchKernelGGL(k_step_advance, dim3(1), dim3(1), 0, ctx->stream, h->step_dev);
    FV_LAUNCH_CHECK(ctx);
    h->step = s + 1;
    return FV_OK;
}

// ---------------------------------------------------------------- graphs
// A steady step enqueues exactly the work of the step before it: graph mode
// (launches sized by capacities, device counts), the ring wake at its full
// n_keep + 1 rows, a conversion every step, particles present, a step that
// sheds (not the last), and no closure or population control (whose sizes
// are decided on the host).  Steady steps alternate between the two wake
// buffers, so two graphs (one per parity) replay them.
static bool hybrid_steady(const fv_hybrid *h)
{
    const fv_hybrid_desc &d = h->d;
    return h->graph_mode == 1 && h->nrow == h->R && d.n_keep < h->R && h->vpm->np_upper > 0 &&
           h->step >= 1 && h->step + 1 < d.num_steps && !h->vpm->closure && h->pop_scheme == 0;
}

static HybridHost host_of(const fv_hybrid *h)
{
    const fv_vpm *f = h->vpm;
    return HybridHost{h->step, h->nrow, h->par, h->seg_pairs, f->np_upper, f->np_host,
                      f->state_epoch, f->np_exact};
}

static void set_host(fv_hybrid *h, const HybridHost &x)
{
    fv_vpm *f = h->vpm;
    h->step = x.step;
    h->nrow = x.nrow;
    h->par = x.par;
    h->seg_pairs = x.seg_pairs;
    f->np_upper = x.np_upper;
    f->np_host = x.np_host;
    f->state_epoch = x.state_epoch;
    f->np_exact = x.np_exact;
}

static HybridSig sig_of(const fv_hybrid *h)
{
    HybridSig g;
    memset(&g, 0, sizeof g);
    g.free_wake = h->d.free_wake;
    g.stretch = h->d.stretch;
    g.precision = h->d.precision;
    g.pop_scheme = h->pop_scheme;
    g.ring_free = h->ring_free;
    g.reuse_on = h->reuse_on;
    g.kernel = h->vpm->kernel;
    g.vprec = h->vpm->precision;
    g.pair_rule = h->vpm->pair_rule;
    g.trunc_x = h->trunc_x;
    g.rlxf = h->vpm->rlxf;
    g.nu = h->vpm->nu;
    g.closure = h->vpm->closure;
    return g;
}

static void batch_drop_graphs(fv_hybrid_batch *b)
{
    for (int p = 0; p < 2; ++p)
        if (b->exec[p]) {
            (void)hipGraphExecDestroy(b->exec[p]);
            b->exec[p] = nullptr;
        }
}

// Captures one step of every run into the graph of wake parity par: each
// run's stream forks from the master stream, enqueues its step and joins
// back.  The host code of the step runs once; its host-side effects are
// recorded (after) and undone (before), and every replay applies them again.
static int batch_capture(fv_hybrid_batch *b, int par)
{
    const size_t n = b->runs.size();
    fv_ctx *ctx0 = b->runs[0]->ctx;
    b->before[par].resize(n);
    b->after[par].resize(n);
    b->sig[par].resize(n);
    std::vector<char> timing(n);
    for (size_t i = 0; i < n; ++i) {
        fv_hybrid *h = b->runs[i];
        b->before[par][i] = host_of(h);
        b->sig[par][i] = sig_of(h);
        timing[i] = h->ctx->timing;
        h->ctx->timing = false;   // no event records inside the graph
    }
    hipGraph_t g = nullptr;
    int rc = FV_OK;
    hipError_t e = hipStreamBeginCapture(b->master, hipStreamCaptureModeThreadLocal);
    const bool began = e == hipSuccess;
    if (began) e = hipEventRecord(b->ev_fork, b->master);
    for (size_t i = 0; i < n && e == hipSuccess && rc == FV_OK; ++i) {
        fv_hybrid *h = b->runs[i];
        e = hipStreamWaitEvent(h->ctx->stream, b->ev_fork, 0);
        if (e == hipSuccess) rc = hybrid_step_one(h);
        if (e == hipSuccess && rc == FV_OK) e = hipEventRecord(b->ev_join[i], h->ctx->stream);
        if (e == hipSuccess && rc == FV_OK) e = hipStreamWaitEvent(b->master, b->ev_join[i], 0);
    }
    const hipError_t e_end = began ? hipStreamEndCapture(b->master, &g) : hipSuccess;
    for (size_t i = 0; i < n; ++i) {
        fv_hybrid *h = b->runs[i];
        b->after[par][i] = host_of(h);
        set_host(h, b->before[par][i]);
        h->ctx->timing = timing[i];
    }
    if (rc == FV_OK && (e != hipSuccess || e_end != hipSuccess))
        rc = set_error(ctx0, FV_E_HIP, "capturing the hybrid step failed: %s",
                       hipGetErrorString(e != hipSuccess ? e : e_end));
    if (rc == FV_OK) {
        e = hipGraphInstantiate(&b->exec[par], g, nullptr, nullptr, 0);
        if (e != hipSuccess) {
            b->exec[par] = nullptr;
            rc = set_error(ctx0, FV_E_HIP, "hipGraphInstantiate of the hybrid step failed: %s",
                           hipGetErrorString(e));
        }
    }
    if (g) (void)hipGraphDestroy(g);
    (void)hipGetLastError();
    return rc;
}

// The master stream's graph launches follow the runs' queued work (join) and
// precede what the runs enqueue afterwards (release).
static int batch_join(fv_hybrid_batch *b)
{
    if (b->on_master) return FV_OK;
    fv_ctx *ctx0 = b->runs[0]->ctx;
    for (size_t i = 0; i < b->runs.size(); ++i) {
        FV_HIP_CHECK(ctx0, hipEventRecord(b->ev_join[i], b->runs[i]->ctx->stream));
        FV_HIP_CHECK(ctx0, hipStreamWaitEvent(b->master, b->ev_join[i], 0));
    }
    b->on_master = true;
    return FV_OK;
}

static int batch_release(fv_hybrid_batch *b)
{
    if (!b->on_master) return FV_OK;
    fv_ctx *ctx0 = b->runs[0]->ctx;
    FV_HIP_CHECK(ctx0, hipEventRecord(b->ev_fork, b->master));
    for (fv_hybrid *h : b->runs) FV_HIP_CHECK(ctx0, hipStreamWaitEvent(h->ctx->stream, b->ev_fork, 0));
    b->on_master = false;
    return FV_OK;
}

static int batch_step(fv_hybrid_batch *b, int64_t nsteps)
{
    for (int64_t k = 0; k < nsteps; ++k) {
        const int par = b->runs[0]->par;
        bool ready = true;
        for (fv_hybrid *h : b->runs) ready = ready && hybrid_steady(h) && h->warm && h->par == par;
        if (ready && b->exec[par]) {
            for (size_t i = 0; i < b->runs.size() && ready; ++i)
                ready = sig_of(b->runs[i]) == b->sig[par][i];
            if (!ready) {
                // a setting changed: the graphs are stale, and the next eager
                // steady step sizes the buffers again
                batch_drop_graphs(b);
                for (fv_hybrid *h : b->runs) h->warm = false;
            }
        }
        int rc = FV_OK;
        if (ready) {
            if (!b->exec[par]) {
                rc = batch_capture(b, par);
                if (rc) return rc;
            }
            rc = batch_join(b);
            if (rc) return rc;
            FV_HIP_CHECK(b->runs[0]->ctx, hipGraphLaunch(b->exec[par], b->master));
            for (size_t i = 0; i < b->runs.size(); ++i) {
                fv_hybrid *h = b->runs[i];
                const HybridHost &x0 = b->before[par][i], &x1 = b->after[par][i];
                HybridHost x = host_of(h);
                x.step += x1.step - x0.step;
                x.nrow = x1.nrow;
                x.par = x1.par;
                x.seg_pairs += x1.seg_pairs - x0.seg_pairs;
                x.np_upper += x1.np_upper - x0.np_upper;
                if (x.np_upper > h->vpm->cap) x.np_upper = h->vpm->cap;
                x.state_epoch += x1.state_epoch - x0.state_epoch;
                x.np_exact = x1.np_exact;
                set_host(h, x);
            }
            ++b->graph_launches;
        } else {
            rc = batch_release(b);
            if (rc) return rc;
            for (fv_hybrid *h : b->runs) {
                const bool steady = hybrid_steady(h);
                rc = hybrid_step_one(h);
                if (rc) return rc;
                // an eager steady step has sized every buffer a capture uses
                h->warm = steady;
            }
            ++b->eager_steps;
        }
    }
    return batch_release(b);
}

static void batch_free(fv_hybrid_batch *b)
{
    if (!b) return;
    if (b->master) (void)hipStreamSynchronize(b->master);
    batch_drop_graphs(b);
    if (b->ev_fork) (void)hipEventDestroy(b->ev_fork);
    for (hipEvent_t e : b->ev_join)
        if (e) (void)hipEventDestroy(e);
    if (b->master) (void)hipStreamDestroy(b->master);
    delete b;
}

static int batch_make(fv_hybrid *const *runs, int32_t n, fv_hybrid_batch **out)
{
    *out = nullptr;
    fv_ctx *ctx0 = runs[0]->ctx;
    fv_hybrid_batch *b = new (std::nothrow) fv_hybrid_batch();
    if (!b) return set_error(ctx0, FV_E_HIP, "out of host memory");
    b->runs.assign(runs, runs + n);
    b->ev_join.assign((size_t)n, nullptr);
    hipError_t e = hipSetDevice(ctx0->device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->master, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->ev_fork, hipEventDisableTiming);
    for (int32_t i = 0; i < n && e == hipSuccess; ++i)
        e = hipEventCreateWithFlags(&b->ev_join[i], hipEventDisableTiming);
    if (e != hipSuccess) {
        batch_free(b);
        return set_error(ctx0, FV_E_HIP, "hybrid batch set-up failed: %s", hipGetErrorString(e));
    }
    *out = b;
    return FV_OK;
}

}  // namespace fv

using namespace fv;

extern "C" {

int fv_hybrid_create(fv_ctx *ctx, const fv_hybrid_desc *desc, fv_hybrid **out)
{
    if (!ctx || !desc || !out) return FV_E_ARG;
    *out = nullptr;
    const fv_hybrid_desc &d = *desc;
    FV_REQUIRE(ctx, d.n_surfaces >= 1 && d.n_surfaces <= 64, "n_surfaces %d outside 1..64",
               d.n_surfaces);
    FV_REQUIRE(ctx, d.nc && d.ns && d.grids, "null surface/grid arrays");
    FV_REQUIRE(ctx, d.num_steps >= 1, "num_steps must be >= 1");
    FV_REQUIRE(ctx, d.dt > 0 && d.rho > 0 && d.sref > 0, "dt, rho and sref must be positive");
    FV_REQUIRE(ctx, d.n_keep >= 0, "n_keep must be >= 0");
    FV_REQUIRE(ctx, d.max_particles >= 1, "max_particles must be >= 1");
    FV_REQUIRE(ctx, d.kernel == FV_KERNEL_GAUSSIAN_ERF || d.kernel == FV_KERNEL_WINCKELMANS,
               "unknown kernel %d", d.kernel);
    FV_REQUIRE(ctx, d.precision == FV_PREC_FP64 || d.precision == FV_PREC_MIXED,
               "unknown precision %d", d.precision);
    const double V = sqrt(d.vinf[0] * d.vinf[0] + d.vinf[1] * d.vinf[1] + d.vinf[2] * d.vinf[2]);
    FV_REQUIRE(ctx, V > 1e-10, "freestream speed must be positive");
    fv_hybrid *h = new (std::nothrow) fv_hybrid();
    if (!h) return set_error(ctx, FV_E_HIP, "out of host memory");
    h->ctx = ctx;
    h->d = d;
    h->d.nc = nullptr;
    h->d.ns = nullptr;
    h->d.grids = nullptr;
    h->R = d.n_keep + 1 < d.num_steps ? d.n_keep + 1 : d.num_steps;
    int64_t P = 0, TV = 0, WV = 0, WR = 0, NS = 0;
    for (int k = 0; k < d.n_surfaces; ++k) {
        if (d.nc[k] < 1 || d.ns[k] < 1) {
            delete h;
            return set_error(ctx, FV_E_ARG, "surface %d has nc=%d ns=%d", k, d.nc[k], d.ns[k]);
        }
        fv_surface sf;
        sf.nc = d.nc[k];
        sf.ns = d.ns[k];
        sf.panel_off = P;
        sf.vert_off = TV;
        sf.wvert_off = WV;
        sf.wring_off = WR;
        h->surf.push_back(sf);
        P += (int64_t)sf.nc * sf.ns;
        TV += (int64_t)(sf.nc + 1) * (sf.ns + 1);
        WV += (h->R + 1) * (sf.ns + 1);
        WR += h->R * sf.ns;
        NS += sf.ns;
    }
    h->P = P;
    h->TV = TV;
    h->WV = WV;
    h->WR = WR;
    h->NS = NS;
    const size_t S = (size_t)d.n_surfaces;
    // doubles: grids + panel geometry + legs + solve + sums + wake + corners + conversion + history
    size_t n = 0;
    const size_t gsz = (size_t)d.num_steps * TV * 3;
    n += gsz;
    n += 4 * 3 * P + 3 * P + 3 * P + P + 3 * P;      // rings, colloc, normal, area, vm_col
    n += 3 * 12 * P;                                 // leg_c, leg_v, vm_leg
    n += (size_t)P * P + 3 * P;                      // aic, rhs, gam, gam_prev
    n += 2 * (size_t)P * P;                          // factor reuse: key, factors
    n += 3 * P + 3 * P + 12 * P + 3 * P;             // vw, vp, vl, force
    n += 2 * (3 * WV + 2 * WR);                      // wake double buffers
    n += 2 * 3 * WV;                                 // free-wake points, velocities
    n += 4 * 3 * WR + 3 * WR + P;                    // ring corners, rg, rage, rrc0, brc0
    n += 4 * 3 * NS + NS;                            // conversion corners + cg
    n += 2 * (size_t)d.num_steps;                    // cl, cdi
    const size_t tab_bytes = sizeof(SegRec) * (size_t)(4 * P + 4 * WR + 4 * P);
    const size_t meta_bytes =
        sizeof(int64_t) * (2 * S + 1) + sizeof(int32_t) * (1 + 2 * S + 2 + P) + 64;
    const size_t bytes = n * sizeof(double) + tab_bytes + meta_bytes + 1024;
    hipError_t e = hipMalloc(&h->block, bytes);
    if (e != hipSuccess) {
        delete h;
        return set_error(ctx, FV_E_HIP, "hipMalloc(%zu) for hybrid case failed: %s", bytes,
                         hipGetErrorString(e));
    }
    char *base = static_cast<char *>(h->block);
    SegRec *tabs = reinterpret_cast<SegRec *>(base);
    h->unit_tab = tabs;
    h->tab = tabs + 4 * P;
    double *p = reinterpret_cast<double *>(base + tab_bytes);
    auto take = [&](size_t k) {
        double *r = p;
        p += k;
        return r;
    };
    h->grids = take(gsz);
    h->br = take(3 * P); h->fr = take(3 * P); h->fl = take(3 * P); h->bl = take(3 * P);
    h->colloc = take(3 * P); h->normal = take(3 * P); h->area = take(P); h->vm_col = take(3 * P);
    h->leg_c = take(12 * P); h->leg_v = take(12 * P); h->vm_leg = take(12 * P);
    h->aic = take((size_t)P * P); h->rhs = take(P); h->gam = take(P); h->gam_prev = take(P);
    h->reuse.key = take((size_t)P * P);
    h->reuse.fac = take((size_t)P * P);
    h->vw = take(3 * P); h->vp = take(3 * P); h->vl = take(12 * P); h->force = take(3 * P);
    for (int b = 0; b < 2; ++b) {
        h->wv[b] = take(3 * WV);
        h->wg[b] = take(WR);
        h->wa[b] = take(WR);
    }
    h->wpts = take(3 * WV); h->wvel = take(3 * WV);
    h->rbr = take(3 * WR); h->rfr = take(3 * WR); h->rfl = take(3 * WR); h->rbl = take(3 * WR);
    h->rg = take(WR); h->rage = take(WR); h->rrc0 = take(WR); h->brc0 = take(P);
    h->cbr = take(3 * NS); h->cfr = take(3 * NS); h->cfl = take(3 * NS); h->cbl = take(3 * NS);
    h->cg = take(NS);
    h->cl = take(d.num_steps); h->cdi = take(d.num_steps);
    // metadata: int64 panel offsets [S], wake ring offsets [S], step counter,
    // then int32 status, nc [S], ns [S], reuse state, reuse count, row order [P]
    int64_t *m64 = reinterpret_cast<int64_t *>(p);
    int32_t *m32p = reinterpret_cast<int32_t *>(m64 + 2 * S + 1);
    h->d_poff = m64;
    h->d_woff = m64 + S;
    h->step_dev = m64 + 2 * S;
    h->info = m32p;
    h->d_nc = m32p + 1;
    h->d_ns = m32p + 1 + S;
    h->reuse.state = m32p + 1 + 2 * S;
    h->reuse.hits = m32p + 2 + 2 * S;
    h->reuse.order = m32p + 3 + 2 * S;
    std::vector<int64_t> po(2 * S + 1, 0);   // panel offsets, wake offsets, step 0
    std::vector<int32_t> m32(1 + 2 * S + 2, 0);   // status, nc, ns, reuse state + hits
    for (size_t k = 0; k < S; ++k) {
        po[k] = h->surf[k].panel_off;
        po[S + k] = h->surf[k].wring_off;
        m32[1 + k] = h->surf[k].nc;
        m32[1 + S + k] = h->surf[k].ns;
    }
    e = hipMemcpy(m64, po.data(), sizeof(int64_t) * po.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemcpy(m32p, m32.data(), sizeof(int32_t) * m32.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemcpy(h->grids, d.grids, sizeof(double) * gsz, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        (void)hipFree(h->block);
        delete h;
        return set_error(ctx, FV_E_HIP, "hybrid setup copy failed: %s", hipGetErrorString(e));
    }
    // fixed per-case arrays
    int rc = fill(ctx, h->rrc0, WR, d.rc0_wake);
    if (!rc) rc = fill(ctx, h->brc0, P, d.rc0_bound);
    if (!rc) rc = fill(ctx, h->cl, d.num_steps, 0.0);
    if (!rc) rc = fill(ctx, h->cdi, d.num_steps, 0.0);
    if (!rc) rc = fill(ctx, h->gam_prev, P, 0.0);
    if (rc) {
        (void)hipFree(h->block);
        delete h;
        return rc;
    }
    rc = fv_vpm_create(ctx, d.max_particles, d.vpm_nu, d.rlxf, d.kernel, d.precision, &h->vpm);
    if (rc) {
        (void)hipFree(h->block);
        delete h;
        return rc;
    }
    // the side stream (default priority) and its two events; without them the
    // solve stays on the context's stream
    if (hipStreamCreateWithFlags(&h->side, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&h->ev_rhs, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&h->ev_solved, hipEventDisableTiming) != hipSuccess) {
        (void)hipGetLastError();
        if (h->side) (void)hipStreamDestroy(h->side);
        if (h->ev_rhs) (void)hipEventDestroy(h->ev_rhs);
        h->side = nullptr;
        h->ev_rhs = h->ev_solved = nullptr;
    }
    FV_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    *out = h;
    return FV_OK;
}

void fv_hybrid_destroy(fv_hybrid *h)
{
    if (!h) return;
    (void)hipStreamSynchronize(h->ctx->stream);
    if (h->side) {
        (void)hipStreamSynchronize(h->side);
        (void)hipStreamDestroy(h->side);
    }
    if (h->ev_rhs) (void)hipEventDestroy(h->ev_rhs);
    if (h->ev_solved) (void)hipEventDestroy(h->ev_solved);
    batch_free(h->self);
    fv_vpm_destroy(h->vpm);
    (void)hipFree(h->block);
    delete h;
}

int fv_hybrid_step(fv_hybrid *h, int64_t nsteps)
{
    if (!h) return FV_E_ARG;
    fv_ctx *ctx = h->ctx;
    FV_REQUIRE(ctx, nsteps >= 0, "negative step count");
    FV_REQUIRE(ctx, h->step + nsteps <= h->d.num_steps, "only %lld of %lld steps remain",
               (long long)(h->d.num_steps - h->step), (long long)h->d.num_steps);
    int32_t ovf = 0;
    FV_HIP_CHECK(ctx, hipMemcpyAsync(&ovf, h->vpm->overflow_dev, sizeof ovf,
                                     hipMemcpyDeviceToHost, ctx->stream));
    FV_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    if (ovf)
        return set_error(ctx, FV_E_CAP, "Particle field would overflow (capacity %lld)",
                         (long long)h->vpm->cap);
    if (h->graph_mode == 1) {
        if (!h->self) {
            int rc = batch_make(&h, 1, &h->self);
            if (rc) return rc;
        }
        return batch_step(h->self, nsteps);
    }
    for (int64_t k = 0; k < nsteps; ++k) {
        int rc = hybrid_step_one(h);
        if (rc) return rc;
    }
    return FV_OK;
}

int fv_hybrid_set_graph(fv_hybrid *h, int32_t on)
{
    if (!h) return FV_E_ARG;
    fv_ctx *ctx = h->ctx;
    FV_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    batch_free(h->self);
    h->self = nullptr;
    FV_REQUIRE(ctx, on >= 0 && on <= 2, "graph mode %d is not 0, 1 or 2", on);
    h->graph_mode = on;
    h->warm = false;
    ctx->fixed_plans = on != 0;
    h->vpm->bound_cap = on != 0;
    return FV_OK;
}

int fv_hybrid_graph_stats(fv_hybrid *h, int64_t *graph_launches, int64_t *eager_steps)
{
    if (!h || !graph_launches || !eager_steps) return FV_E_ARG;
    *graph_launches = h->self ? h->self->graph_launches : 0;
    *eager_steps = h->self ? h->self->eager_steps : 0;
    return FV_OK;
}

int fv_hybrid_batch_create(fv_hybrid *const *runs, int32_t n, fv_hybrid_batch **out)
{
    if (!out) return FV_E_ARG;
    *out = nullptr;
    if (!runs || n < 1 || !runs[0]) return FV_E_ARG;
    fv_ctx *ctx0 = runs[0]->ctx;
    for (int32_t i = 0; i < n; ++i) {
        FV_REQUIRE(ctx0, runs[i] != nullptr, "run %d is null", i);
        FV_REQUIRE(ctx0, runs[i]->ctx->device == ctx0->device, "runs on different devices");
        for (int32_t j = 0; j < i; ++j)
            FV_REQUIRE(ctx0, runs[i] != runs[j], "run %d appears twice", i);
    }
    return batch_make(runs, n, out);
}

int fv_hybrid_batch_step(fv_hybrid_batch *b, int64_t nsteps)
{
    if (!b) return FV_E_ARG;
    fv_ctx *ctx0 = b->runs[0]->ctx;
    FV_REQUIRE(ctx0, nsteps >= 0, "negative step count");
    for (fv_hybrid *h : b->runs) {
        FV_REQUIRE(ctx0, h->step + nsteps <= h->d.num_steps, "a run has only %lld steps left",
                   (long long)(h->d.num_steps - h->step));
        int32_t ovf = 0;
        FV_HIP_CHECK(ctx0, hipMemcpyAsync(&ovf, h->vpm->overflow_dev, sizeof ovf,
                                          hipMemcpyDeviceToHost, h->ctx->stream));
        FV_HIP_CHECK(ctx0, hipStreamSynchronize(h->ctx->stream));
        if (ovf)
            return set_error(ctx0, FV_E_CAP, "Particle field would overflow (capacity %lld)",
                             (long long)h->vpm->cap);
    }
    return batch_step(b, nsteps);
}

int fv_hybrid_batch_stats(fv_hybrid_batch *b, int64_t *graph_launches, int64_t *eager_steps)
{
    if (!b || !graph_launches || !eager_steps) return FV_E_ARG;
    *graph_launches = b->graph_launches;
    *eager_steps = b->eager_steps;
    return FV_OK;
}

void fv_hybrid_batch_destroy(fv_hybrid_batch *b) { batch_free(b); }

int fv_hybrid_current_step(fv_hybrid *h, int64_t *step)
{
    if (!h || !step) return FV_E_ARG;
    *step = h->step;
    return FV_OK;
}

int fv_hybrid_read(fv_hybrid *h, double *cl, double *cdi, double *gamma)
{
    if (!h) return FV_E_ARG;
    fv_ctx *ctx = h->ctx;
    if (cl)
        FV_HIP_CHECK(ctx, hipMemcpyAsync(cl, h->cl, sizeof(double) * h->d.num_steps,
                                         hipMemcpyDeviceToHost, ctx->stream));
    if (cdi)
        FV_HIP_CHECK(ctx, hipMemcpyAsync(cdi, h->cdi, sizeof(double) * h->d.num_steps,
                                         hipMemcpyDeviceToHost, ctx->stream));
    if (gamma)
        FV_HIP_CHECK(ctx, hipMemcpyAsync(gamma, h->gam, sizeof(double) * h->P,
                                         hipMemcpyDeviceToHost, ctx->stream));
    FV_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    int32_t ovf = 0;
    FV_HIP_CHECK(ctx, hipMemcpy(&ovf, h->vpm->overflow_dev, sizeof ovf, hipMemcpyDeviceToHost));
    if (ovf)
        return set_error(ctx, FV_E_CAP, "Particle field would overflow (capacity %lld)",
                         (long long)h->vpm->cap);
    return FV_OK;
}

fv_vpm *fv_hybrid_particles(fv_hybrid *h) { return h ? h->vpm : nullptr; }

int fv_hybrid_set_flags(fv_hybrid *h, int32_t free_wake, int32_t stretch)
{
    if (!h) return FV_E_ARG;
    h->d.free_wake = free_wake ? 1 : 0;
    h->d.stretch = stretch ? 1 : 0;
    return FV_OK;
}

int fv_hybrid_set_population(fv_hybrid *h, int32_t scheme, int64_t max_particles, double eps,
                             int64_t protect, double protect_dist)
{
    if (!h) return FV_E_ARG;
    FV_REQUIRE(h->ctx, scheme >= 0 && scheme <= 2, "unknown population scheme %d", scheme);
    FV_REQUIRE(h->ctx, max_particles >= 0 && eps >= 0.0 && protect >= 0 && protect_dist >= 0.0,
               "negative population parameter");
    h->pop_scheme = scheme;
    h->pop_max = max_particles;
    h->pop_eps = eps;
    h->pop_protect = protect;
    h->pop_protect_dist = protect_dist;
    return FV_OK;
}

int fv_hybrid_set_ring_wake(fv_hybrid *h, int32_t free_wake, double truncate_x)
{
    if (!h) return FV_E_ARG;
    FV_REQUIRE(h->ctx, !std::isnan(truncate_x), "truncation x is NaN");
    h->ring_free = free_wake ? 1 : 0;
    h->trunc_x = truncate_x;
    return FV_OK;
}

int fv_hybrid_ledger_host(fv_hybrid *h, double *out3)
{
    if (!h || !out3) return FV_E_ARG;
    fv_ctx *ctx = h->ctx;
    // the stored wake rings as corners, then particles + rings
    int64_t W = 0;
    const int cur = h->par;
    for (auto &sf : h->surf) {
        if (h->nrow > 0) {
            hipLaunchKernelGGL(k_wake_corners, dim3((unsigned)ceil_div(h->nrow * sf.ns, 128)),
                               dim3(128), 0, ctx->stream, h->wv[cur] + 3 * sf.wvert_off,
                               h->wg[cur] + sf.wring_off, h->wa[cur] + sf.wring_off, sf.ns,
                               (int64_t)0, h->nrow, W, h->rbr, h->rfr, h->rfl, h->rbl, h->rg,
                               h->rage);
            FV_LAUNCH_CHECK(ctx);
        }
        W += h->nrow * sf.ns;
    }
    int rc = vpm_ledger(h->vpm, W, h->rbr, h->rfr, h->rfl, h->rbl, h->rg, h->vw);
    if (rc) return rc;
    FV_HIP_CHECK(ctx, hipMemcpyAsync(out3, h->vw, 3 * sizeof(double), hipMemcpyDeviceToHost,
                                     ctx->stream));
    FV_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return FV_OK;
}

int fv_hybrid_set_precision(fv_hybrid *h, int32_t precision)
{
    if (!h) return FV_E_ARG;
    FV_REQUIRE(h->ctx, precision == FV_PREC_FP64 || precision == FV_PREC_MIXED,
               "unknown precision %d", precision);
    if (precision != h->d.precision) {
        // the stored factors are of the other precision's type
        FV_HIP_CHECK(h->ctx, hipMemsetAsync(h->reuse.state, 0, sizeof(int32_t), h->ctx->stream));
    }
    h->d.precision = precision;
    h->vpm->precision = precision;
    return FV_OK;
}

int fv_hybrid_set_solve_reuse(fv_hybrid *h, int32_t on)
{
    if (!h) return FV_E_ARG;
    if (!on) FV_HIP_CHECK(h->ctx, hipMemsetAsync(h->reuse.state, 0, sizeof(int32_t), h->ctx->stream));
    h->reuse_on = on ? 1 : 0;
    return FV_OK;
}

int fv_hybrid_solve_reuse_count(fv_hybrid *h, int64_t *reused)
{
    if (!h) return FV_E_ARG;
    fv_ctx *ctx = h->ctx;
    FV_REQUIRE(ctx, reused, "null output");
    int32_t v = 0;
    FV_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    if (h->side) FV_HIP_CHECK(ctx, hipStreamSynchronize(h->side));
    FV_HIP_CHECK(ctx, hipMemcpy(&v, h->reuse.hits, sizeof v, hipMemcpyDeviceToHost));
    *reused = v;
    return FV_OK;
}

int fv_hybrid_pair_counts(fv_hybrid *h, double *seg_pairs, double *particle_pairs)
{
    if (!h) return FV_E_ARG;
    if (seg_pairs) *seg_pairs = h->seg_pairs;
    if (particle_pairs) {
        fv_ctx *ctx = h->ctx;
        double dev[2] = {0.0, 0.0};
        FV_HIP_CHECK(ctx, hipMemcpyAsync(dev, h->vpm->pairs_dev, sizeof dev,
                                         hipMemcpyDeviceToHost, ctx->stream));
        FV_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
        *particle_pairs = dev[0] + dev[1] + h->vpm->particle_pairs;
    }
    return FV_OK;
}

}  // extern "C"
