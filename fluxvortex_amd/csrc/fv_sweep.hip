// fv_sweep.hip -- batched aeroelastic hybrid UVLM instances (a flutter-speed
// sweep) on one GPU.
//
// Reference: tests/benchmark_goland.py:100-243 loops over airspeeds, each a
// PteraSoftware UVLM run coupled to BeamFE by AeroelasticSolver
// (src/fluxvortex/aeroelastic_solver.py:63-221, beam_fe.py:231-351).  Here
// every airspeed is an instance of the hybrid N_keep step and all instances
// advance together, two launches per step.  The first spreads the pair work
// of every instance over the whole device, one workgroup per unit of three
// roles:
//   sw_geom      AIC (unit rings . normals at the collocation points) and the
//                unit-strength bound-ring velocities at the leg centres of
//                the step's lattice, one lane per entry;
//   sw_wake      wake-ring velocities at the collocation points, the leg
//                centres (and free wake vertices), one lane per target of a
//                segment chunk, partial sums per chunk;
//   sw_induce    particle -> collocation-point velocities, one workgroup per
//                (instance, 1024-particle chunk), partial sums per chunk; it
//                also applies the previous step's prescribed-wake convection
//                (x += V dt) to the particles it stages;
// and one workgroup per instance then runs the serial rest of the step with
// its arrays in LDS:
//   k_sw_step    rings / frames / legs, right-hand side, LU solve (factor
//                reuse when the AIC is unchanged), leg velocities, Kutta-
//                Joukowski loads, CL / CDi, station loads -> beam nodes, one
//                Newmark step, the next lattice, wake shedding, N_keep
//                conversion.
// The launches are captured once in a hipGraph and replayed per step (every
// per-step quantity lives in device memory).
//
// The UVLM statement is the device driver's (fv_hybrid.hip) and the oracle's
// (oracle/fv_oracle_uvlm.c); the coupling follows AeroelasticSolver: station
// lift = lift_sign * F_z summed over the strip (the reference's -F_z,
// aeroelastic_solver.py:135-139), station moment = sum of x_colloc * lift, the
// panel force sharing each line between two rows half and half.  With
// `feedback` the deformed lattice carries the next step's aerodynamics (every
// vertex moved with the heave / twist at its own |y|); without it the
// aerodynamics stay on the undeformed lattice, as in the reference, whose
// vertex mutation does not reach its solver (its stored Goland histories
// carry loads independent of the deformation).  Particles are a prescribed
// wake moved by V dt per step.
#include <cmath>
#include <cstring>
#include <new>
#include <vector>

#include "fv_lu.h"

namespace fv {

constexpr int kSwMaxNc = 16, kSwMaxNs = 32, kSwMaxP = 64, kSwMaxNodes = 17;
constexpr int kSwMaxFree = 3 * (kSwMaxNodes - 1);
constexpr int kSwMaxL = 3 * kSwMaxP + kSwMaxNs;   // right / front / left legs + trailing-edge back legs
constexpr int kSwBlock = 256;       // induction workgroup
constexpr int kSwStepBlock = 1024;  // step workgroup: 16 waves, one instance per CU
constexpr int kSwGeomBlock = 256;   // geometry workgroup: one AIC / leg-influence entry per lane
constexpr int kSwChunk = 1024;   // particles per induction partial
constexpr int kSwStreams = 4;    // source streams per induction wavefront (targets <= 64 / streams)
constexpr int kSwWChunkSmall = 16;    // wake segments per partial sum (wakes up to 64 chunks)
constexpr int kSwWChunkLarge = 256;   // wake segments per partial sum (longer wakes)

struct SwInst {
    int nc, ns, P, nv, R, n_keep, kernel, couple, n_nodes, nfree;
    int status;                  // 0; k + 1: zero pivot at step k; -2: particle overflow
    int has_prev, par, nrow, has_wprev, prec, feedback;
    int wchunk;                  // wake segments per sw_wake partial sum (<= kSwBlock)
    int shift;                   // the particles still owe this step's x += V dt
    int free_wake;               // free ring wake: vertex rows move with V + u
    int truncate;                // rings whose corner-mean x exceeds trunc_x lose Gamma
    int64_t num_steps, step, cap, np, max_chunks, max_wchunks, max_targets;
    double V[3], rho, dt, nu, rc0b, rc0w, sigma, sref, x_ea, relax, beta, gamma, lift_sign;
    double trunc_x;
    // offsets (doubles) into the state block
    int64_t o_grid0, o_coln, o_colp, o_legp, o_gamp, o_gam, o_wv[2], o_wg[2], o_wa[2];
    int64_t o_pos, o_al, o_sg, o_part, o_M, o_C, o_K, o_d, o_v, o_a;
    int64_t o_wdef, o_tdef, o_wprev, o_tprev, o_yn, o_cl, o_cdi, o_tw, o_tt;
    int64_t o_legn, o_wpart;     // the step's leg centres; wake partial sums
    int64_t o_ten, o_vvel;       // the step's trailing-edge points; free-vertex velocities
    int64_t o_wsum;              // the wake partial sums added up, 3 per target
    int64_t o_gcur;              // the step's lattice (3 nv)
    int64_t o_aic, o_bleg;       // sw_geom: AIC (P x P), leg influences (NL x P x 3)
    // factor reuse of the circulation solve: the AIC last factored, its
    // factors, the composed row order (as doubles) and a state word (1.0:
    // valid); a step whose AIC equals it bit for bit (a lattice that does
    // not move) skips the factorization, with the same solution bit for bit
    int64_t o_lkey, o_lfac, o_lord, o_lst;
};

// Free ring wake: vertex rows 0 .. nfree - 1 move with the induced velocity
// (row 0: the trailing edge of the step's lattice; rows >= 1: the wake's old
// vertex rows; rows beyond n_keep hold no circulation).
__device__ __forceinline__ int sw_nfree(const SwInst &I)
{
    return I.free_wake ? (I.nrow < I.n_keep ? I.nrow : I.n_keep) + 1 : 0;
}

// ------------------------------------------------------------ geometry helpers
__device__ __forceinline__ void sw_lerp(const double *a, const double *b, double t, double *o)
{
    o[0] = a[0] + t * (b[0] - a[0]);
    o[1] = a[1] + t * (b[1] - a[1]);
    o[2] = a[2] + t * (b[2] - a[2]);
}

__device__ __forceinline__ const double *sw_v(const double *g, int ns, int i, int j)
{
    return g + 3 * (i * (ns + 1) + j);
}

// Ring of panel (i, j): front legs at the quarter chord, back legs at the next
// panel's quarter chord (last row: one panel length behind the front leg).
__device__ void sw_ring(const double *g, int nc, int ns, int i, int j, double *Br, double *Fr,
                        double *Fl, double *Bl)
{
    const double *v00 = sw_v(g, ns, i, j), *v01 = sw_v(g, ns, i, j + 1);
    const double *v10 = sw_v(g, ns, i + 1, j), *v11 = sw_v(g, ns, i + 1, j + 1);
    sw_lerp(v00, v10, 0.25, Fl);
    sw_lerp(v01, v11, 0.25, Fr);
    if (i < nc - 1) {
        sw_lerp(v10, sw_v(g, ns, i + 2, j), 0.25, Bl);
        sw_lerp(v11, sw_v(g, ns, i + 2, j + 1), 0.25, Br);
    } else {
        for (int c = 0; c < 3; ++c) {
            Bl[c] = Fl[c] + (v10[c] - v00[c]);
            Br[c] = Fr[c] + (v11[c] - v01[c]);
        }
    }
}

// Collocation point (3/4 chord, mid span), unit normal (Fr - Bl) x (Fl - Br), area.
__device__ void sw_frame(const double *g, int ns, int i, int j, double *col, double *nrm,
                         double *area)
{
    const double *v00 = sw_v(g, ns, i, j), *v01 = sw_v(g, ns, i, j + 1);
    const double *v10 = sw_v(g, ns, i + 1, j), *v11 = sw_v(g, ns, i + 1, j + 1);
    double a[3], b[3];
    sw_lerp(v00, v10, 0.75, a);
    sw_lerp(v01, v11, 0.75, b);
    for (int c = 0; c < 3; ++c) col[c] = 0.5 * (a[c] + b[c]);
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

// numpy.interp: linear inside [xp0, xp_last], the end values outside.
__device__ double sw_interp(double x, const double *xp, const double *fp, int n)
{
    if (x <= xp[0]) return fp[0];
    if (x >= xp[n - 1]) return fp[n - 1];
    int j = 0;
    while (j + 1 < n - 1 && xp[j + 1] <= x) ++j;
    if (x == xp[j]) return fp[j];
    const double slope = (fp[j + 1] - fp[j]) / (xp[j + 1] - xp[j]);
    return slope * (x - xp[j]) + fp[j];
}

// Vertex deformed by heave w and twist theta at its |y| (rotation about x_ea):
// z += w + (x - x_ea) sin(theta), x += (x - x_ea)(cos(theta) - 1)
// (aeroelastic_solver.py:176-198).
__device__ void sw_deform(const double *b, const double *yn, const double *wn, const double *tn,
                          int nn, double x_ea, double *o)
{
    const double y = fabs(b[1]);
    const double w = sw_interp(y, yn, wn, nn), th = sw_interp(y, yn, tn, nn);
    const double xr = b[0] - x_ea;
    o[0] = b[0] + xr * (cos(th) - 1.0);
    o[1] = b[1];
    o[2] = b[2] + (w + xr * sin(th));
}

// Segment record of leg k (Br->Fr, Fr->Fl, Fl->Bl, Bl->Br) of a ring with
// strength G, core rc0 and age (warp_kernels.py:46-96 constants).
__device__ SegRec sw_seg(const double *s, const double *t, double G, double rc0, double age,
                         double nu)
{
    SegRec r;
    r.sx = s[0]; r.sy = s[1]; r.sz = s[2];
    r.ex = t[0]; r.ey = t[1]; r.ez = t[2];
    const double r0x = t[0] - s[0], r0y = t[1] - s[1], r0z = t[2] - s[2];
    const double r0sq = r0x * r0x + r0y * r0y + r0z * r0z;
    const double r0len = sqrt(r0sq);
    const double rcsq = rc0 * rc0 + kFourLamb * (nu + kSquire * fabs(G)) * age;
    r.c1 = G * kInv4Pi;
    r.c2 = r0sq * rcsq;
    r.tol = (r0len < 2.22e-16) ? INFINITY : r0len * 1.0e-10;
    r.pad = 0.0;
    return r;
}

// Segment pair in the instance's precision (mixed: FP32 pair term, FP64 sums).
__device__ __forceinline__ void sw_pair(int prec, const double *x, const SegRec &sr, double *u)
{
    if (prec == FV_PREC_MIXED)
        seg_pair_mixed(x[0], x[1], x[2], sr, u[0], u[1], u[2]);
    else
        seg_pair_f64(x[0], x[1], x[2], sr, u[0], u[1], u[2]);
}

// Leg centres of panel p (slots 3 p + l for the right / front / left legs, 3 P
// + j for the trailing-edge back leg of (nc - 1, j)), as k_sw_step forms them.
__device__ void sw_leg_centres(const double *g, int nc, int ns, int P, int p, double *lc)
{
    const int i = p / ns, j = p % ns;
    double R[12];
    sw_ring(g, nc, ns, i, j, R, R + 3, R + 6, R + 9);
    for (int l = 0; l < 4; ++l) {
        if (l == 3 && i != nc - 1) continue;
        const int q = l < 3 ? 3 * p + l : 3 * P + j;
        const double *a = R + 3 * l, *b = R + 3 * ((l + 1) & 3);
        for (int c = 0; c < 3; ++c) lc[3 * q + c] = 0.5 * (a[c] + b[c]);
    }
}

// ----------------------------------------------------------- particle induction
// Lane l of a wavefront takes target l % P of source stream l / P (streams =
// min(kSwStreams, 64 / P) per wavefront, so P <= 32 keeps every lane busy and
// streams x P <= 256); the streams' sums are added in stream order.  The particles of the chunk are
// staged once, with the previous step's convection x += V dt applied and
// written back when the instance owes it.
template <int KERNEL, int PREC>
__device__ void sw_induce_body(const SwInst &I, double *__restrict__ S, double *part,
                               int64_t chunk, PartRec *tile, double *tgt, double *red)
{
    const int t = threadIdx.x, P = I.P;
    const int G = (64 / P < kSwStreams) ? 64 / P : kSwStreams;
    const int lane = t & 63, tp = lane % P, gs = lane / P;
    const bool act = gs < G;
    const int stream = (t >> 6) * G + gs, nstream = (kSwBlock / 64) * G;
    const int64_t c0 = chunk * kSwChunk;
    const int64_t c1 = (c0 + kSwChunk < I.np) ? c0 + kSwChunk : I.np;
    for (int e = t; e < 3 * P; e += kSwBlock) tgt[e] = S[I.o_coln + e];
    using Op = PartOp<KERNEL, PREC, 0>;
    if constexpr (KERNEL == 0) gauss_stage();
    typename Op::Tgt tg;
    typename Op::Acc acc;
    Op::init(acc);
    __syncthreads();
    tg.x = tgt[3 * tp]; tg.y = tgt[3 * tp + 1]; tg.z = tgt[3 * tp + 2];
    tg.ax = tg.ay = tg.az = 0.0;
    double *pos = S + I.o_pos;
    const double *al = S + I.o_al, *sg = S + I.o_sg;
    const bool shift = I.shift != 0;
    const double ox = I.V[0] * I.dt, oy = I.V[1] * I.dt, oz = I.V[2] * I.dt;
    for (int64_t b0 = c0; b0 < c1; b0 += kSwBlock) {
        const int cnt = (int)((c1 - b0) < kSwBlock ? (c1 - b0) : kSwBlock);
        __syncthreads();
        if (t < cnt) {
            const int64_t j = b0 + t;
            double x = pos[3 * j], y = pos[3 * j + 1], z = pos[3 * j + 2];
            if (shift) {
                x += ox;
                y += oy;
                z += oz;
                pos[3 * j] = x;
                pos[3 * j + 1] = y;
                pos[3 * j + 2] = z;
            }
            tile[t] = make_part_rec<PREC>(KERNEL, x, y, z, al[3 * j], al[3 * j + 1], al[3 * j + 2],
                                          sg[j]);
        }
        __syncthreads();
        if (act)
            for (int k = stream; k < cnt; k += nstream) Op::pair(tg, tile[k], acc);
    }
    Op::finish(tg, acc);
    if (act)
        for (int c = 0; c < 3; ++c) red[(stream * P + tp) * 3 + c] = acc.v[c];
    __syncthreads();
    if (t < P)
        for (int c = 0; c < 3; ++c) {
            double s = 0.0;
            for (int g = 0; g < nstream; ++g) s += red[(g * P + t) * 3 + c];
            part[(chunk * P + t) * 3 + c] = s;
        }
}

// LDS of the induction role: a source tile, the targets, the stream sums.
constexpr size_t kSwIndLds = sizeof(PartRec) * kSwBlock + sizeof(double) * (3 * kSwMaxP +
                                                                           (kSwBlock / 64) * kSwMaxP * 3);

__device__ void sw_induce(const SwInst &I, double *__restrict__ S, int64_t chunk, unsigned char *lds)
{
    if (I.status != 0 || I.step >= I.num_steps) return;
    if (chunk * kSwChunk >= I.np) return;   // uniform over the workgroup
    PartRec *tile = reinterpret_cast<PartRec *>(lds);
    double *tgt = reinterpret_cast<double *>(lds + sizeof(PartRec) * kSwBlock);
    double *red = tgt + 3 * kSwMaxP;   // streams x P <= 4 x 64
    double *part = S + I.o_part;
    const bool mixed = I.prec == FV_PREC_MIXED;
    if (I.kernel == FV_KERNEL_WINCKELMANS) {
        if (mixed) sw_induce_body<1, 1>(I, S, part, chunk, tile, tgt, red);
        else sw_induce_body<1, 0>(I, S, part, chunk, tile, tgt, red);
    } else {
        if (mixed) sw_induce_body<0, 1>(I, S, part, chunk, tile, tgt, red);
        else sw_induce_body<0, 0>(I, S, part, chunk, tile, tgt, red);
    }
}

// ------------------------------------------------------------------ ring wake
// Wake-ring velocities at the T = P + NL + NV targets of the step (collocation
// points, the leg centres in slot order, then with a free ring wake the NV =
// nfree (ns + 1) vertices of rows 0 .. nfree - 1) from the segments [c0, c1)
// of one chunk (segment k: leg k & 3 of ring k >> 2; at most kSwBlock, staged
// in LDS at once), one target per lane of a group of kSwBlock targets, partial
// sums wpart[(chunk T + target) 3 + c].
constexpr size_t kSwWakeLds = sizeof(SegRec) * kSwBlock;

__device__ void sw_wake(const SwInst &I, double *__restrict__ S, int64_t chunk, int tgroup,
                        unsigned char *lds)
{
    if (I.status != 0 || I.step >= I.num_steps) return;
    SegRec *tile = reinterpret_cast<SegRec *>(lds);
    const int ns = I.ns, P = I.P, NL = 3 * P + ns;
    const int T = P + NL + sw_nfree(I) * (ns + 1);
    const int64_t nseg = 4 * (int64_t)I.nrow * ns;
    const int64_t c0 = chunk * I.wchunk;
    const int tz = tgroup * kSwBlock;
    if (c0 >= nseg || tz >= T) return;   // uniform over the workgroup
    const int cnt = (int)((nseg - c0 < I.wchunk) ? nseg - c0 : I.wchunk);   // <= kSwBlock
    const double *wv = S + I.o_wv[I.par], *wg = S + I.o_wg[I.par], *wa = S + I.o_wa[I.par];
    const int t = threadIdx.x;
    if (t < cnt) {
        const int k = (int)(c0 + t);   // < 4 nrow ns < 2^31
        const int rg = k >> 2;
        const int leg = k & 3, r = rg / ns, j = rg % ns;
        const double *c4[4] = {wv + 3 * ((r + 1) * (ns + 1) + j + 1),   // Br
                               wv + 3 * (r * (ns + 1) + j + 1),         // Fr
                               wv + 3 * (r * (ns + 1) + j),             // Fl
                               wv + 3 * ((r + 1) * (ns + 1) + j)};      // Bl
        tile[t] = sw_seg(c4[leg], c4[(leg + 1) & 3], wg[rg], I.rc0w, wa[rg], I.nu);
    }
    const int tq = tz + t;
    const bool act = tq < T;
    const int tc = act ? tq : 0;
    const int tv = tc - P - NL;   // free vertex: row tv / (ns + 1), column tv % (ns + 1)
    const double *x = tc < P ? S + I.o_coln + 3 * tc
                    : tv < 0 ? S + I.o_legn + 3 * (tc - P)
                    : tv < ns + 1 ? S + I.o_ten + 3 * tv
                                  : wv + 3 * tv;
    const double xp[3] = {x[0], x[1], x[2]};
    double u[3] = {0.0, 0.0, 0.0};
    __syncthreads();
    if (!act) return;
    const int prec = I.prec;
#pragma unroll 2
    for (int k = 0; k < cnt; ++k) sw_pair(prec, xp, tile[k], u);
    double *part = S + I.o_wpart + chunk * T * 3;
    for (int c = 0; c < 3; ++c) part[3 * tq + c] = u[c];
}

// ------------------------------------------------------------------ the step
// LDS plan (doubles) of k_sw_step at the limits above.
constexpr int kSwNv = (kSwMaxNc + 1) * (kSwMaxNs + 1);
constexpr int kSwLdaMax = kSwMaxP | 1;
constexpr int kL_g = 0;                                   // grid 3 nv
constexpr int kL_ring = kL_g + 3 * kSwNv;                 // Br Fr Fl Bl 12 P
constexpr int kL_col = kL_ring + 12 * kSwMaxP;            // collocation 3 P
constexpr int kL_nrm = kL_col + 3 * kSwMaxP;              // normal 3 P
constexpr int kL_area = kL_nrm + 3 * kSwMaxP;             // area P
constexpr int kL_vmc = kL_area + kSwMaxP;                 // motion velocity 3 P
constexpr int kL_lc = kL_vmc + 3 * kSwMaxP;               // leg centres 3 NL
constexpr int kL_lv = kL_lc + 3 * kSwMaxL;                // leg vectors 3 NL
constexpr int kL_vml = kL_lv + 3 * kSwMaxL;               // leg motion velocity 3 NL
constexpr int kL_aic = kL_vml + 3 * kSwMaxL;              // AIC P x lda
constexpr int kL_rhs = kL_aic + kSwMaxP * kSwLdaMax;      // rhs / gamma P
constexpr int kL_rdg = kL_rhs + kSwMaxP;                  // 1 / diag P
constexpr int kL_perm = kL_rdg + kSwMaxP;                 // pivots, row order (int) 2 P
constexpr int kL_vl = kL_perm + kSwMaxP;                  // leg velocities 3 NL
constexpr int kL_F = kL_vl + 3 * kSwMaxL;                 // panel forces 3 P
constexpr int kL_PF = kL_F + 3 * kSwMaxP;                 // panel shares, front-leg forces 6 P
constexpr int kL_st = kL_PF + 6 * kSwMaxP;                // station lift, moment, node, xi 4 ns
constexpr int kL_mck = kL_st + 4 * kSwMaxNs;              // Newmark M, C, K^-1 3 nf^2
constexpr int kL_dva = kL_mck + 3 * kSwMaxFree * kSwMaxFree;   // Newmark d, v, a 3 nf
constexpr int kL_beam = kL_dva + 3 * kSwMaxFree;          // nodal forces / vectors
constexpr int kL_end = kL_beam + 3 * kSwMaxNodes + 4 * kSwMaxFree + 4 * kSwMaxNodes + 8;
constexpr size_t kSwLdsBytes = sizeof(double) * kL_end;
static_assert(kSwLdsBytes + 2048 <= 160 * 1024, "k_sw_step LDS plan exceeds the 160 KB of a CU");

// ------------------------------------------------------------- step geometry
// The step's lattice (S[o_gcur]) -> entry e < P^2: the AIC, A[i][j] = (ring
// j's four unit-strength legs at collocation point i) . n_i, in the instance's
// precision; entry P^2 + q P + j: the velocity of ring j's unit legs at leg
// centre q, stored at o_bleg + (j NL + q) 3 (k_sw_step contracts it with the
// step's circulations; the bound rings have age 0, so their cores do not
// depend on the circulation and the velocity is linear in it).
constexpr size_t kSwGeomLds = sizeof(double) * (3 * kSwNv + 12 * kSwMaxP + 6 * kSwMaxP + 3 * kSwMaxL);

__device__ void sw_geom(const SwInst &I, double *__restrict__ S, int chunk, unsigned char *lds)
{
    if (I.status != 0 || I.step >= I.num_steps) return;
    const int nc = I.nc, ns = I.ns, P = I.P, nv = I.nv, NL = 3 * P + ns;
    const int E = P * P + NL * P;
    if (chunk * kSwGeomBlock >= E) return;   // uniform over the workgroup
    double *g = reinterpret_cast<double *>(lds), *ring = g + 3 * kSwNv;
    double *col = ring + 12 * kSwMaxP, *nrm = col + 3 * kSwMaxP, *lc = nrm + 3 * kSwMaxP;
    const int t = threadIdx.x;
    for (int v = t; v < 3 * nv; v += kSwGeomBlock) g[v] = S[I.o_gcur + v];
    __syncthreads();
    for (int p = t; p < P; p += kSwGeomBlock) {
        const int i = p / ns, j = p % ns;
        double *Br = ring + 12 * p;
        sw_ring(g, nc, ns, i, j, Br, Br + 3, Br + 6, Br + 9);
        double ar;
        sw_frame(g, ns, i, j, col + 3 * p, nrm + 3 * p, &ar);
        sw_leg_centres(g, nc, ns, P, p, lc);
    }
    __syncthreads();
    const int e = chunk * kSwGeomBlock + t;
    if (e >= E) return;
    const bool aic = e < P * P;
    const int f = aic ? e : e - P * P;
    const int q = f / P, j = f % P;   // AIC: row (collocation point) q; else leg centre q
    const double *x = aic ? col + 3 * q : lc + 3 * q;
    const double *R = ring + 12 * j;
    double u[3] = {0.0, 0.0, 0.0};
    for (int k = 0; k < 4; ++k) {
        const SegRec sr = sw_seg(R + 3 * k, R + 3 * ((k + 1) & 3), 1.0, I.rc0b, 0.0, I.nu);
        sw_pair(I.prec, x, sr, u);
    }
    if (aic) {
        S[I.o_aic + e] = u[0] * nrm[3 * q] + u[1] * nrm[3 * q + 1] + u[2] * nrm[3 * q + 2];
    } else {
        double *b = S + I.o_bleg + ((int64_t)j * NL + q) * 3;
        for (int c = 0; c < 3; ++c) b[c] = u[c];
    }
}

// The three device-wide roles of a step in one launch (one workgroup = one
// unit of one role, heaviest first): particle induction (n x gi units: the
// instance's 1024-particle chunks), wake induction (n x gw x tg units: segment
// chunk x group of kSwBlock targets), geometry (n x gg units of kSwGeomBlock
// entries).  The roles write disjoint outputs.
constexpr size_t kSwPreLds = kSwIndLds > kSwGeomLds ? (kSwIndLds > kSwWakeLds ? kSwIndLds : kSwWakeLds)
                                                    : (kSwGeomLds > kSwWakeLds ? kSwGeomLds : kSwWakeLds);
static_assert(kSwGeomBlock == kSwBlock, "the roles share one workgroup size");

__global__ __launch_bounds__(kSwBlock) void k_sw_pre(const SwInst *__restrict__ insts,
                                                     double *__restrict__ S, int n, int gi, int gw,
                                                     int tg, int gg)
{
    __shared__ __attribute__((aligned(16))) unsigned char lds[kSwPreLds];
    int64_t b = blockIdx.x;
    const int64_t ni = (int64_t)n * gi, nw = (int64_t)n * gw * tg;
    if (b < ni) {
        sw_induce(insts[b / gi], S, b % gi, lds);
        return;
    }
    b -= ni;
    if (b < nw) {
        const int64_t per = (int64_t)gw * tg, r = b % per;
        sw_wake(insts[b / per], S, r / tg, (int)(r % tg), lds);
        return;
    }
    b -= nw;
    sw_geom(insts[b / gg], S, (int)(b % gg), lds);
}

__global__ __launch_bounds__(kSwStepBlock) void k_sw_step(SwInst *__restrict__ insts,
                                                      double *__restrict__ S)
{
    extern __shared__ __attribute__((aligned(16))) double L[];
    __shared__ SwInst I;
    __shared__ int fail_s;
    const int t = threadIdx.x;
    if (t == 0) {
        I = insts[blockIdx.x];
        fail_s = 0;
    }
    __syncthreads();
    if (I.status != 0 || I.step >= I.num_steps) return;
    const int nc = I.nc, ns = I.ns, P = I.P, nv = I.nv, nn = I.n_nodes, nf = I.nfree;
    const int64_t s = I.step;
    const double dt = I.dt, *V = I.V;
    double *g = L + kL_g, *ring = L + kL_ring, *col = L + kL_col, *nrm = L + kL_nrm;
    double *area = L + kL_area, *vmc = L + kL_vmc, *lc = L + kL_lc, *lv = L + kL_lv;
    double *vml = L + kL_vml, *A = L + kL_aic, *rhs = L + kL_rhs, *rdg = L + kL_rdg;
    int *perm = reinterpret_cast<int *>(L + kL_perm);
    double *vl = L + kL_vl, *F = L + kL_F;
    double *PF = L + kL_PF, *Ff = PF + 3 * kSwMaxP;
    double *Mb = L + kL_mck, *Cb = Mb + nf * nf, *Kb = Cb + nf * nf;
    double *db = L + kL_dva, *vb = db + nf, *ab = vb + nf;
    // the wake of this step, read in place
    const int nrow = I.nrow, cur = I.par;
    const double *wv = S + I.o_wv[cur], *wg = S + I.o_wg[cur], *wa = S + I.o_wa[cur];
    double *bm = L + kL_beam;
    const double *grid0 = S + I.o_grid0, *yn = S + I.o_yn;
    double *wdef = S + I.o_wdef, *tdef = S + I.o_tdef;
    const int lda = P | 1;
    const int NL = 3 * P + ns;   // leg slots: 3 p + l, then 3 P + j for the back leg of (nc-1, j)
    const bool couple = I.couple && s >= 1 && s <= I.num_steps - 2;
    // ---- lattice of this step (written by the previous step or k_sw_init)
    // and the Newmark matrices / state, staged in LDS
    for (int v = t; v < 3 * nv; v += kSwStepBlock) g[v] = S[I.o_gcur + v];
    if (couple) {
        for (int e = t; e < nf * nf; e += kSwStepBlock) {
            Mb[e] = S[I.o_M + e];
            Cb[e] = S[I.o_C + e];
            Kb[e] = S[I.o_K + e];
        }
        for (int r = t; r < nf; r += kSwStepBlock) {
            db[r] = S[I.o_d + r];
            vb[r] = S[I.o_v + r];
            ab[r] = S[I.o_a + r];
        }
    }
    __syncthreads();
    for (int p = t; p < P; p += kSwStepBlock) {
        const int i = p / ns, j = p % ns;
        double *Br = ring + 12 * p, *Fr = Br + 3, *Fl = Br + 6, *Bl = Br + 9;
        sw_ring(g, nc, ns, i, j, Br, Fr, Fl, Bl);
        sw_frame(g, ns, i, j, col + 3 * p, nrm + 3 * p, area + p);
        const double *cp = S + I.o_colp + 3 * p;
        for (int c = 0; c < 3; ++c) vmc[3 * p + c] = I.has_prev ? (col[3 * p + c] - cp[c]) / dt : 0.0;
        for (int l = 0; l < 4; ++l) {
            // right, front, left legs; the back leg only on the trailing edge
            if (l == 3 && i != nc - 1) continue;
            const int q = l < 3 ? 3 * p + l : 3 * P + j;
            const double *a = Br + 3 * l, *b = Br + 3 * ((l + 1) & 3);
            const double *lp = S + I.o_legp + 3 * q;
            for (int c = 0; c < 3; ++c) {
                const double cc = 0.5 * (a[c] + b[c]);
                lc[3 * q + c] = cc;
                lv[3 * q + c] = b[c] - a[c];
                vml[3 * q + c] = I.has_prev ? (cc - lp[c]) / dt : 0.0;
            }
        }
    }
    // ---- wake-ring velocities at the step's targets: sw_wake's chunk
    // partials added per (target, component), four interleaved running sums
    // combined in a fixed order
    const int W = nrow * ns;
    const int nfree = sw_nfree(I), NV = nfree * (ns + 1);
    const int T = P + NL + NV;
    const int nwch = (int)((4 * (int64_t)W + I.wchunk - 1) / I.wchunk);
    for (int e = t; e < 3 * T; e += kSwStepBlock) {
        const double *pp = S + I.o_wpart + e;
        double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
        int ch = 0;
        for (; ch + 4 <= nwch; ch += 4) {
            a0 += pp[(int64_t)ch * 3 * T];
            a1 += pp[(int64_t)(ch + 1) * 3 * T];
            a2 += pp[(int64_t)(ch + 2) * 3 * T];
            a3 += pp[(int64_t)(ch + 3) * 3 * T];
        }
        for (; ch < nwch; ++ch) a0 += pp[(int64_t)ch * 3 * T];
        S[I.o_wsum + e] = (a0 + a1) + (a2 + a3);
    }
    __syncthreads();
    // ---- right-hand side
    const int nchunk = (int)((I.np + kSwChunk - 1) / kSwChunk);
    for (int p = t; p < P; p += kSwStepBlock) {
        double w3[3], q3[3] = {0.0, 0.0, 0.0};
        for (int c = 0; c < 3; ++c) w3[c] = S[I.o_wsum + 3 * p + c];
        for (int ch = 0; ch < nchunk; ++ch)
            for (int c = 0; c < 3; ++c) q3[c] += S[I.o_part + (ch * P + p) * 3 + c];
        const double *n = nrm + 3 * p;
        const double fs = (V[0] - vmc[3 * p]) * n[0] + (V[1] - vmc[3 * p + 1]) * n[1] +
                          (V[2] - vmc[3 * p + 2]) * n[2];
        const double wn = w3[0] * n[0] + w3[1] * n[1] + w3[2] * n[2];
        const double qn = I.np > 0 ? q3[0] * n[0] + q3[1] * n[1] + q3[2] * n[2] : 0.0;
        rhs[p] = -((wn + qn) + fs);
    }
    __syncthreads();
    // ---- circulation solve (LU with partial pivoting in LDS) of sw_geom's
    // AIC; the factorization is skipped when the AIC equals the last factored
    // one
    double *lkey = S + I.o_lkey, *lfac = S + I.o_lfac, *lord = S + I.o_lord;
    int *order = perm + P;   // the pivots composed into one gather order (kL_perm holds 2 P ints)
    const bool check = S[I.o_lst] == 1.0;
    const bool differ = mat_move<kSwStepBlock, 2>(S + I.o_aic, P, A, lda, P,
                                                  check ? (const double *)lkey : (const double *)nullptr);
    if (!__syncthreads_or((!check || differ) ? 1 : 0)) {
        mat_move<kSwStepBlock, 2>(lfac, P, A, lda, P, (const double *)nullptr);
        for (int i = t; i < P; i += kSwStepBlock) order[i] = (int)lord[i];
    } else {
        mat_move<kSwStepBlock, 2>(A, lda, lkey, P, P, (const double *)nullptr);
        if (t == 0) S[I.o_lst] = 0.0;
        // LDS panels: the step kernel's 1024-lane workgroup leaves 128 VGPRs a lane
        const int inf = lu_blocked<double, double, false>(A, lda, P, perm, (double *)nullptr,
                                                          (double *)nullptr);
        if (inf != 0) {
            if (t == 0) {
                insts[blockIdx.x].status = (int)s + 1;
            }
            return;
        }
        for (int i = t; i < P; i += kSwStepBlock) order[pivot_track(perm, P, i)] = i;
        mat_move<kSwStepBlock, 2>(A, lda, lfac, P, P, (const double *)nullptr);
        __syncthreads();
        for (int i = t; i < P; i += kSwStepBlock) lord[i] = (double)order[i];
        if (t == 0) S[I.o_lst] = 1.0;
    }
    // L y = P rhs, then U x = y (the same substitutions on either path)
    for (int i = t; i < P; i += kSwStepBlock) rdg[i] = rhs[i];
    __syncthreads();
    for (int i = t; i < P; i += kSwStepBlock) rhs[i] = rdg[order[i]];
    __syncthreads();
    wave_fwd_subst<double, double, false>(A, lda, P, rhs);   // LDS form: 128 VGPRs a lane
    __syncthreads();
    for (int i = t; i < P; i += kSwStepBlock) rdg[i] = 1.0 / A[i * lda + i];
    __syncthreads();
    wave_back_subst<double, double>(A, lda, P, rhs, rdg);
    __syncthreads();
    const double *gam = rhs;
    // ---- velocities at the leg centres: the bound rings through sw_geom's
    // unit influences contracted with the circulations (in ring order) + the
    // wake rings
    {
        const double *B = S + I.o_bleg;
        for (int e = t; e < 3 * NL; e += kSwStepBlock) {
            double a = 0.0;
            for (int j = 0; j < P; ++j) a += gam[j] * B[(int64_t)j * 3 * NL + e];
            vl[e] = a + S[I.o_wsum + 3 * P + e];
        }
    }
    __syncthreads();
    // ---- free ring wake: velocities at vertex rows 0 .. nfree - 1 from the
    // bound rings (this step's circulation) + the wake rings (sw_wake's
    // partial sums), stored for the shedding below
    if (NV > 0) {
        double *vv = S + I.o_vvel;
        for (int e = t; e < NV; e += kSwStepBlock) {
            const int r = e / (ns + 1), j = e % (ns + 1);
            double x[3];
            if (r == 0) {
                const double *va = sw_v(g, ns, nc - 1, j), *vb = sw_v(g, ns, nc, j);
                for (int c = 0; c < 3; ++c) x[c] = vb[c] + 0.25 * (vb[c] - va[c]);
            } else {
                for (int c = 0; c < 3; ++c) x[c] = wv[3 * e + c];
            }
            double u[3] = {0.0, 0.0, 0.0};
            for (int k = 0; k < 4 * P; ++k) {
                const int jr = k >> 2, leg = k & 3;
                const double *R = ring + 12 * jr;
                const SegRec sr = sw_seg(R + 3 * leg, R + 3 * ((leg + 1) & 3), gam[jr], I.rc0b, 0.0, I.nu);
                sw_pair(I.prec, x, sr, u);
            }
            for (int c = 0; c < 3; ++c) vv[3 * e + c] = u[c] + S[I.o_wsum + 3 * (P + NL + e) + c];
        }
    }
    // ---- loads: Kutta-Joukowski on each leg with the net circulation of its
    // line (right / left against the spanwise neighbour, front against the ring
    // ahead, the trailing-edge back leg against the newest wake ring) + the
    // unsteady term -rho A dGamma/dt n
    const double *gprev = S + I.o_gamp;
    for (int p = t; p < P; p += kSwStepBlock) {
        const int i = p / ns, j = p % ns;
        const double G = gam[p];
        const double gw = nrow > 0 ? wg[j] : 0.0;
        const double ge[4] = {(j + 1 < ns) ? G - gam[p + 1] : G, (i > 0) ? G - gam[p - ns] : G,
                              (j > 0) ? G - gam[p - 1] : G, (i == nc - 1) ? G - gw : 0.0};
        double Fp[3] = {0.0, 0.0, 0.0}, fl[4][3];
        for (int l = 0; l < 4; ++l) {
            if (l == 3 && i != nc - 1) {
                fl[3][0] = fl[3][1] = fl[3][2] = 0.0;
                continue;
            }
            const int q = l < 3 ? 3 * p + l : 3 * P + j;
            const double u[3] = {V[0] - vml[3 * q] + vl[3 * q], V[1] - vml[3 * q + 1] + vl[3 * q + 1],
                                 V[2] - vml[3 * q + 2] + vl[3 * q + 2]};
            const double *Lg = lv + 3 * q;
            const double f = I.rho * ge[l];
            fl[l][0] = f * (u[1] * Lg[2] - u[2] * Lg[1]);
            fl[l][1] = f * (u[2] * Lg[0] - u[0] * Lg[2]);
            fl[l][2] = f * (u[0] * Lg[1] - u[1] * Lg[0]);
            for (int c = 0; c < 3; ++c) Fp[c] += fl[l][c];
        }
        const double dG = s == 0 ? 0.0 : (G - gprev[p]) / dt;
        const double fu = -I.rho * area[p] * dG;
        for (int c = 0; c < 3; ++c) {
            F[3 * p + c] = Fp[c] + fu * nrm[3 * p + c];
            // the panel's share: the line to the row ahead half and half
            PF[3 * p + c] = fl[0][c] + fl[2][c] + fl[1][c] * (i > 0 ? 0.5 : 1.0) + fl[3][c] +
                            fu * nrm[3 * p + c];
            Ff[3 * p + c] = fl[1][c];
        }
    }
    __syncthreads();
    const double V2 = V[0] * V[0] + V[1] * V[1] + V[2] * V[2], Vn = sqrt(V2);
    const double dd[3] = {V[0] / Vn, V[1] / Vn, V[2] / Vn};
    double l0 = -dd[2], l2 = dd[0];
    const double ln = sqrt(l0 * l0 + l2 * l2);
    l0 /= ln;
    l2 /= ln;
    if (t == 0) {
        double Ft[3] = {0.0, 0.0, 0.0};
        for (int p = 0; p < P; ++p)
            for (int c = 0; c < 3; ++c) Ft[c] += F[3 * p + c];
        const double qs = 0.5 * I.rho * V2 * I.sref;
        S[I.o_cl + s] = (Ft[0] * l0 + Ft[2] * l2) / qs;
        S[I.o_cdi + s] = (Ft[0] * dd[0] + Ft[1] * dd[1] + Ft[2] * dd[2]) / qs;
    }
    // ---- structure: station loads -> beam nodes, Newmark, next deformation
    double *Fn = bm, *Fe = bm + 3 * kSwMaxNodes, *dn = Fe + kSwMaxFree, *an = dn + kSwMaxFree;
    double *tmp = an + kSwMaxFree, *wnew = tmp + kSwMaxFree, *tnew = wnew + kSwMaxNodes;
    if (couple) {
        // per station (one lane each): lift, moment and the node pair it loads
        double *st = L + kL_st;
        for (int j = t; j < ns; j += kSwStepBlock) {
            double lift = 0.0, mom = 0.0;
            for (int i = 0; i < nc; ++i) {
                const int p = i * ns + j;
                const double fz = PF[3 * p + 2] + (i + 1 < nc ? 0.5 * Ff[3 * (p + ns) + 2] : 0.0);
                const double lp = I.lift_sign * fz;
                lift += lp;
                mom += col[3 * p] * lp;
            }
            const double y = col[3 * j + 1];   // first chordwise panel of the station
            double node = -1.0, xi = 0.0;      // node -1: no load (the root station)
            if (fabs(y) > 1e-6) {
                int idx = -1;
                for (int k = 0; k < nn; ++k)
                    if (yn[k] <= y) idx = k;
                idx = idx < 0 ? 0 : (idx > nn - 2 ? nn - 2 : idx);
                const double Le = yn[idx + 1] - yn[idx];
                xi = (y - yn[idx]) / Le;
                xi = xi < 0.0 ? 0.0 : (xi > 1.0 ? 1.0 : xi);
                node = (double)idx;
            }
            st[j] = lift;
            st[kSwMaxNs + j] = mom;
            st[2 * kSwMaxNs + j] = node;
            st[3 * kSwMaxNs + j] = xi;
        }
        __syncthreads();
        // per node (one lane each): its stations' shares added in station order
        for (int k = t; k < nn; k += kSwStepBlock) {
            double fw = 0.0, fm = 0.0;
            for (int j = 0; j < ns; ++j) {
                const double node = st[2 * kSwMaxNs + j];
                if (node < 0.0) continue;
                const int idx = (int)node;
                const double lift = st[j], mom = st[kSwMaxNs + j], xi = st[3 * kSwMaxNs + j];
                if (idx == k) {
                    fw += lift * (1 - xi);
                    fm += mom * (1 - xi);
                } else if (idx + 1 == k) {
                    fw += lift * xi;
                    fm += mom * xi;
                }
            }
            Fn[3 * k] = fw;
            Fn[3 * k + 1] = 0.0;
            Fn[3 * k + 2] = fm;
        }
        __syncthreads();
        const double be = I.beta, ga = I.gamma;
        const double a0 = 1.0 / (be * dt * dt), a1 = 1.0 / (be * dt), a2 = 1.0 / (2.0 * be) - 1.0;
        const double a3 = ga / (be * dt), a4 = ga / be - 1.0, a5 = dt * (ga / (2.0 * be) - 1.0);
        // F_eff = F + M (a0 d + a1 v + a2 a) + C (a3 d + a4 v + a5 a)
        for (int r = t; r < nf; r += kSwStepBlock) {
            double sm = 0.0, sc = 0.0;
            for (int c = 0; c < nf; ++c) {
                sm += Mb[r * nf + c] * (a0 * db[c] + a1 * vb[c] + a2 * ab[c]);
                sc += Cb[r * nf + c] * (a3 * db[c] + a4 * vb[c] + a5 * ab[c]);
            }
            Fe[r] = Fn[3 + r] + sm + sc;
        }
        __syncthreads();
        double *d = S + I.o_d, *v = S + I.o_v, *a = S + I.o_a;
        for (int r = t; r < nf; r += kSwStepBlock) {
            double x = 0.0;
            for (int c = 0; c < nf; ++c) x += Kb[r * nf + c] * Fe[c];
            dn[r] = x;
            an[r] = a0 * (dn[r] - db[r]) - a1 * vb[r] - a2 * ab[r];
            v[r] = vb[r] + dt * ((1.0 - ga) * ab[r] + ga * an[r]);
            a[r] = an[r];
            d[r] = dn[r];
        }
        __syncthreads();
        // nodal heave / twist (one lane per node), relaxed against the last
        double *wp = S + I.o_wprev, *tp = S + I.o_tprev;
        for (int k = t; k < nn; k += kSwStepBlock) {
            double w = k == 0 ? 0.0 : dn[3 * (k - 1)];
            double th = k == 0 ? 0.0 : dn[3 * (k - 1) + 2];
            if (I.has_wprev) {
                w = I.relax * w + (1.0 - I.relax) * wp[k];
                th = I.relax * th + (1.0 - I.relax) * tp[k];
            }
            wp[k] = w;
            tp[k] = th;
            if (k == nn - 1) {
                S[I.o_tw + s] = w;
                S[I.o_tt + s] = th;
            }
            // without feedback the aerodynamics stay on the undeformed lattice
            wnew[k] = I.feedback ? w : 0.0;
            tnew[k] = I.feedback ? th : 0.0;
        }
        __syncthreads();
        if (t == 0) I.has_wprev = 1;
    } else if (t == 0) {
        for (int k = 0; k < nn; ++k) {
            wnew[k] = 0.0;
            tnew[k] = 0.0;
        }
    }
    __syncthreads();
    // ---- wake shedding, conversion, next lattice (not after the last step)
    const bool more = s + 1 < I.num_steps;
    const int nxt = 1 - cur;
    if (more) {
        const int nrow_new = nrow + 1 < I.R ? nrow + 1 : I.R;
        double *wvn = S + I.o_wv[nxt], *wgn = S + I.o_wg[nxt], *wan = S + I.o_wa[nxt];
        const double ox = V[0] * dt, oy = V[1] * dt, oz = V[2] * dt;
        // vertex rows: 0 = trailing edge of the next lattice; 1 = current trailing
        // edge + V dt; r + 1 = old row r + V dt
        for (int e = t; e < (nrow_new + 1) * (ns + 1); e += kSwStepBlock) {
            const int r = e / (ns + 1), j = e % (ns + 1);
            double o[3];
            if (r == 0) {
                double va[3], vb[3];
                sw_deform(grid0 + 3 * ((nc - 1) * (ns + 1) + j), yn, wnew, tnew, nn, I.x_ea, va);
                sw_deform(grid0 + 3 * (nc * (ns + 1) + j), yn, wnew, tnew, nn, I.x_ea, vb);
                for (int c = 0; c < 3; ++c) o[c] = vb[c] + 0.25 * (vb[c] - va[c]);
            } else if (r - 1 < nfree) {
                // free vertex: old row r - 1 moves with V + u (frozen snapshot)
                const double *w = S + I.o_vvel + 3 * ((r - 1) * (ns + 1) + j);
                double x0[3];
                if (r == 1) {
                    const double *va = sw_v(g, ns, nc - 1, j), *vb = sw_v(g, ns, nc, j);
                    for (int c = 0; c < 3; ++c) x0[c] = vb[c] + 0.25 * (vb[c] - va[c]);
                } else {
                    for (int c = 0; c < 3; ++c) x0[c] = wv[3 * ((r - 1) * (ns + 1) + j) + c];
                }
                for (int c = 0; c < 3; ++c) o[c] = x0[c] + (V[c] + w[c]) * dt;
            } else if (r == 1) {
                const double *va = sw_v(g, ns, nc - 1, j), *vb = sw_v(g, ns, nc, j);
                for (int c = 0; c < 3; ++c) o[c] = vb[c] + 0.25 * (vb[c] - va[c]);
                o[0] += ox;
                o[1] += oy;
                o[2] += oz;
            } else {
                const double *w = wv + 3 * ((r - 1) * (ns + 1) + j);
                o[0] = w[0] + ox;
                o[1] = w[1] + oy;
                o[2] = w[2] + oz;
            }
            for (int c = 0; c < 3; ++c) wvn[3 * e + c] = o[c];
        }
        for (int e = t; e < nrow_new * ns; e += kSwStepBlock) {
            const int r = e / ns, j = e % ns;
            if (r == 0) {
                wgn[e] = gam[(nc - 1) * ns + j];
                wan[e] = 0.0;
            } else {
                wgn[e] = wg[(r - 1) * ns + j];
                wan[e] = wa[(r - 1) * ns + j] + dt;
            }
        }
        __syncthreads();
        // truncation by ring corner-mean x (Fr + Fl + Bl + Br) / 4, after shedding
        if (I.truncate) {
            for (int e = t; e < nrow_new * ns; e += kSwStepBlock) {
                const int r = e / ns, j = e % ns;
                const double *a = wvn + 3 * (r * (ns + 1) + j), *b = a + 3;
                const double *c = wvn + 3 * ((r + 1) * (ns + 1) + j), *f = c + 3;
                const double cx = (b[0] + a[0] + c[0] + f[0]) * 0.25;
                if (cx > I.trunc_x) wgn[e] = 0.0;
            }
            __syncthreads();
        }
        // conversion of ring row n_keep (span order, legs Fr->Fl, Fl->Bl, Bl->Br,
        // Br->Fr): wavefront 0, lane = 4 j + leg of a 64-leg chunk; the legs
        // that emit a particle take consecutive slots in (j, leg) order from a
        // ballot prefix count
        if (I.n_keep < nrow_new && t < 64) {
            const int kr = I.n_keep;
            int64_t np = I.np;
            double *pos = S + I.o_pos, *al = S + I.o_al, *sg = S + I.o_sg;
            bool over = false;
            for (int e0 = 0; e0 < 4 * ns; e0 += 64) {
                const int e = e0 + t, j = e >> 2, l = e & 3;
                const bool lane_ok = e < 4 * ns;
                const double G = lane_ok ? wgn[kr * ns + j] : 0.0;
                const bool ring_ok = lane_ok && fabs(G) >= 1e-15;
                double pp[3] = {0.0, 0.0, 0.0}, aa[3] = {0.0, 0.0, 0.0};
                if (ring_ok) {
                    const double *FL = wvn + 3 * (kr * (ns + 1) + j), *FR = FL + 3;
                    const double *BL = wvn + 3 * ((kr + 1) * (ns + 1) + j), *BR = BL + 3;
                    const double *ends[4][2] = {{FR, FL}, {FL, BL}, {BL, BR}, {BR, FR}};
                    for (int c = 0; c < 3; ++c) {
                        pp[c] = 0.5 * (ends[l][0][c] + ends[l][1][c]);
                        aa[c] = G * (ends[l][1][c] - ends[l][0][c]);
                    }
                }
                const bool emit = ring_ok && aa[0] * aa[0] + aa[1] * aa[1] + aa[2] * aa[2] > 1e-30;
                const unsigned long long mask = __ballot(emit);
                const int64_t slot = np + __popcll(mask & ((1ull << t) - 1ull));
                if (emit && slot < I.cap) {
                    for (int c = 0; c < 3; ++c) {
                        pos[3 * slot + c] = pp[c];
                        al[3 * slot + c] = aa[c];
                    }
                    sg[slot] = I.sigma;
                }
                over = over || (emit && slot >= I.cap);
                np += __popcll(mask);
                if (ring_ok && l == 0) wgn[kr * ns + j] = 0.0;
            }
            if (__ballot(over) != 0ull && t == 0) fail_s = 1;
            if (t == 0) I.np = np;
        }
        __syncthreads();
        if (fail_s) {
            if (t == 0) insts[blockIdx.x].status = -2;
            return;
        }
        // the particles owe x += V dt (prescribed wake): sw_induce applies it
        // when it stages them for the next step
        // next lattice: deformation, collocation points and leg centres for the
        // next step's launches
        for (int k = t; k < nn; k += kSwStepBlock) {
            wdef[k] = wnew[k];
            tdef[k] = tnew[k];
        }
        double *gn = g;   // the current lattice is no longer needed
        __syncthreads();
        for (int v = t; v < nv; v += kSwStepBlock) {
            sw_deform(grid0 + 3 * v, yn, wnew, tnew, nn, I.x_ea, gn + 3 * v);
            for (int c = 0; c < 3; ++c) S[I.o_gcur + 3 * v + c] = gn[3 * v + c];
        }
        __syncthreads();
        for (int p = t; p < P; p += kSwStepBlock) {
            double cn[3], nn3[3], ar;
            sw_frame(gn, ns, p / ns, p % ns, cn, nn3, &ar);
            for (int c = 0; c < 3; ++c) S[I.o_coln + 3 * p + c] = cn[c];
            sw_leg_centres(gn, nc, ns, P, p, S + I.o_legn);
        }
        if (I.free_wake)
            for (int j = t; j <= ns; j += kSwStepBlock) {
                const double *va = sw_v(gn, ns, nc - 1, j), *vb = sw_v(gn, ns, nc, j);
                for (int c = 0; c < 3; ++c)
                    S[I.o_ten + 3 * j + c] = vb[c] + 0.25 * (vb[c] - va[c]);
            }
        if (t == 0) {
            I.nrow = nrow_new;
            I.par = nxt;
        }
    }
    // ---- carry this step's state
    for (int p = t; p < P; p += kSwStepBlock) {
        for (int c = 0; c < 3; ++c) S[I.o_colp + 3 * p + c] = col[3 * p + c];
        for (int c = 0; c < 9; ++c) S[I.o_legp + 9 * p + c] = lc[9 * p + c];
        if (p / ns == nc - 1)
            for (int c = 0; c < 3; ++c) S[I.o_legp + 3 * (3 * P + p % ns) + c] = lc[3 * (3 * P + p % ns) + c];
        S[I.o_gamp + p] = gam[p];
        S[I.o_gam + p] = gam[p];
    }
    __syncthreads();
    if (t == 0) {
        SwInst &o = insts[blockIdx.x];
        o.np = I.np;
        o.nrow = I.nrow;
        o.par = I.par;
        o.has_wprev = I.has_wprev;
        o.has_prev = 1;
        o.shift = more ? 1 : 0;
        o.step = s + 1;
    }
}

// The undeformed lattice (the first step's, deformed by the zero initial
// heave / twist exactly as every later step's), its collocation points, leg
// centres and trailing-edge points: the first step's launch inputs.
__global__ __launch_bounds__(64) void k_sw_init(const SwInst *__restrict__ insts,
                                                double *__restrict__ S)
{
    const SwInst &I = insts[blockIdx.x];
    double *g = S + I.o_gcur;
    for (int v = threadIdx.x; v < I.nv; v += 64)
        sw_deform(S + I.o_grid0 + 3 * v, S + I.o_yn, S + I.o_wdef, S + I.o_tdef, I.n_nodes, I.x_ea,
                  g + 3 * v);
    __syncthreads();
    for (int p = threadIdx.x; p < I.P; p += 64) {
        double cn[3], nn3[3], ar;
        sw_frame(g, I.ns, p / I.ns, p % I.ns, cn, nn3, &ar);
        for (int c = 0; c < 3; ++c) S[I.o_coln + 3 * p + c] = cn[c];
        sw_leg_centres(g, I.nc, I.ns, I.P, p, S + I.o_legn);
    }
    if (I.free_wake)
        for (int j = threadIdx.x; j <= I.ns; j += 64) {
            const double *va = sw_v(g, I.ns, I.nc - 1, j);
            const double *vb = sw_v(g, I.ns, I.nc, j);
            for (int c = 0; c < 3; ++c) S[I.o_ten + 3 * j + c] = vb[c] + 0.25 * (vb[c] - va[c]);
        }
}

}  // namespace fv

using namespace fv;

struct fv_sweep {
    fv_ctx *ctx = nullptr;
    int n = 0;
    std::vector<SwInst> host;
    SwInst *dinst = nullptr;
    double *block = nullptr;
    int64_t max_chunks = 1, max_wchunks = 1, max_tchunks = 1, max_gchunks = 1, max_steps = 0;
    int64_t done_steps = 0;
    int use_graph = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

static int sweep_launch(fv_sweep *sw)
{
    fv_ctx *ctx = sw->ctx;
    const int64_t blocks = (int64_t)sw->n * (sw->max_chunks + sw->max_wchunks * sw->max_tchunks +
                                             sw->max_gchunks);
    hipLaunchKernelGGL(k_sw_pre, dim3((unsigned)blocks), dim3(kSwBlock), 0, ctx->stream, sw->dinst,
                       sw->block, sw->n, (int)sw->max_chunks, (int)sw->max_wchunks,
                       (int)sw->max_tchunks, (int)sw->max_gchunks);
    FV_LAUNCH_CHECK(ctx);
    hipLaunchKernelGGL(k_sw_step, dim3((unsigned)sw->n), dim3(kSwStepBlock), kSwLdsBytes, ctx->stream,
                       sw->dinst, sw->block);
    FV_LAUNCH_CHECK(ctx);
    return FV_OK;
}

extern "C" {

int fv_sweep_create(fv_ctx *ctx, const fv_sweep_desc *descs, int32_t n, int32_t use_graph,
                    fv_sweep **out)
{
    if (!ctx || !descs || !out) return FV_E_ARG;
    *out = nullptr;
    FV_REQUIRE(ctx, n >= 1 && n <= 65535, "instance count %d outside 1..65535", n);
    fv_sweep *sw = new (std::nothrow) fv_sweep();
    if (!sw) return set_error(ctx, FV_E_HIP, "out of host memory");
    sw->ctx = ctx;
    sw->n = n;
    sw->use_graph = use_graph ? 1 : 0;
    int64_t off = 0;
    auto take = [&](int64_t k) { const int64_t o = off; off += (k + 3) & ~int64_t(3); return o; };
    for (int i = 0; i < n; ++i) {
        const fv_sweep_desc &d = descs[i];
        const double V = sqrt(d.vinf[0] * d.vinf[0] + d.vinf[1] * d.vinf[1] + d.vinf[2] * d.vinf[2]);
        const int nfree = 3 * (d.n_nodes - 1);
        if (!(d.nc >= 1 && d.nc <= kSwMaxNc && d.ns >= 1 && d.ns <= kSwMaxNs &&
              d.nc * d.ns <= kSwMaxP && d.grid && d.num_steps >= 1 && d.dt > 0 && d.rho > 0 &&
              d.sref > 0 && V > 1e-10 && d.n_keep >= 0 &&
              d.max_particles >= 1 && d.n_nodes >= 2 && d.n_nodes <= kSwMaxNodes && d.y_nodes &&
              (d.kernel == FV_KERNEL_GAUSSIAN_ERF || d.kernel == FV_KERNEL_WINCKELMANS) &&
              (d.precision == FV_PREC_FP64 || d.precision == FV_PREC_MIXED) &&
              (!d.couple || (d.M && d.C && d.Keff_inv && d.d0 && d.v0 && d.a0)))) {
            delete sw;
            return set_error(ctx, FV_E_ARG, "sweep instance %d: invalid or out-of-limit description",
                             i);
        }
        SwInst I{};
        I.nc = d.nc;
        I.ns = d.ns;
        I.P = d.nc * d.ns;
        I.nv = (d.nc + 1) * (d.ns + 1);
        const int64_t R = (int64_t)d.n_keep + 1 < d.num_steps ? (int64_t)d.n_keep + 1 : d.num_steps;
        if (R > (int64_t)1 << 24) {
            delete sw;
            return set_error(ctx, FV_E_ARG, "sweep instance %d: %lld kept wake rows exceed the limit",
                             i, (long long)R);
        }
        I.R = (int)R;
        I.free_wake = d.free_ring_wake ? 1 : 0;
        I.truncate = d.truncate_wake ? 1 : 0;
        I.trunc_x = d.truncate_x;
        I.n_keep = d.n_keep;
        I.kernel = d.kernel;
        I.prec = d.precision;
        I.couple = d.couple ? 1 : 0;
        I.n_nodes = d.n_nodes;
        I.nfree = nfree;
        I.num_steps = d.num_steps;
        I.cap = d.max_particles;
        for (int c = 0; c < 3; ++c) I.V[c] = d.vinf[c];
        I.rho = d.rho;
        I.dt = d.dt;
        I.nu = d.nu;
        I.rc0b = d.rc0_bound;
        I.rc0w = d.rc0_wake;
        I.sigma = d.sigma;
        I.sref = d.sref;
        I.x_ea = d.x_ea;
        I.relax = d.relax;
        I.feedback = d.feedback ? 1 : 0;
        I.lift_sign = d.lift_sign;
        I.beta = d.beta;
        I.gamma = d.gamma;
        I.max_chunks = ceil_div(d.max_particles, kSwChunk);
        if (I.max_chunks > sw->max_chunks) sw->max_chunks = I.max_chunks;
        if (d.num_steps > sw->max_steps) sw->max_steps = d.num_steps;
        I.o_grid0 = take(3 * I.nv);
        I.o_coln = take(3 * I.P);
        I.o_colp = take(3 * I.P);
        I.o_legp = take(3 * (3 * I.P + d.ns));
        I.o_gamp = take(I.P);
        I.o_gam = take(I.P);
        for (int b = 0; b < 2; ++b) {
            I.o_wv[b] = take(3 * (I.R + 1) * (d.ns + 1));
            I.o_wg[b] = take(I.R * d.ns);
            I.o_wa[b] = take(I.R * d.ns);
        }
        I.o_pos = take(3 * I.cap);
        I.o_al = take(3 * I.cap);
        I.o_sg = take(I.cap);
        I.o_part = take(3 * I.P * I.max_chunks);
        {
            // the wake holds at most R rows when a step starts; the free vertex
            // rows are at most min(R, n_keep) + 1.  Short wakes take small
            // segment chunks (more workgroups, shorter per-lane loops).
            const int64_t nseg = 4 * (int64_t)I.R * d.ns;
            I.wchunk = ceil_div(nseg, kSwWChunkSmall) <= 64 ? kSwWChunkSmall : kSwWChunkLarge;
            I.max_wchunks = ceil_div(nseg, I.wchunk);
            if (I.max_wchunks < 1) I.max_wchunks = 1;
            if (I.max_wchunks > sw->max_wchunks) sw->max_wchunks = I.max_wchunks;
            const int64_t nv = I.free_wake
                                   ? ((int64_t)(I.R < d.n_keep ? I.R : d.n_keep) + 1) * (d.ns + 1)
                                   : 0;
            I.max_targets = 4 * (int64_t)I.P + d.ns + nv;
            const int64_t tch = ceil_div(I.max_targets, kSwBlock);
            if (tch > sw->max_tchunks) sw->max_tchunks = tch;
            I.o_legn = take(3 * (3 * I.P + d.ns));
            I.o_wpart = take(3 * I.max_targets * I.max_wchunks);
            I.o_wsum = take(3 * I.max_targets);
            I.o_ten = take(3 * (d.ns + 1));
            I.o_vvel = take(3 * (nv > 0 ? nv : 1));
        }
        {
            const int64_t NL = 3 * (int64_t)I.P + d.ns;
            I.o_gcur = take(3 * I.nv);
            I.o_aic = take((int64_t)I.P * I.P);
            I.o_bleg = take(3 * NL * I.P);
            const int64_t gch = ceil_div((int64_t)I.P * I.P + NL * I.P, kSwGeomBlock);
            if (gch > sw->max_gchunks) sw->max_gchunks = gch;
        }
        I.o_M = take((int64_t)nfree * nfree);
        I.o_C = take((int64_t)nfree * nfree);
        I.o_K = take((int64_t)nfree * nfree);
        I.o_d = take(nfree);
        I.o_v = take(nfree);
        I.o_a = take(nfree);
        I.o_wdef = take(d.n_nodes);
        I.o_tdef = take(d.n_nodes);
        I.o_wprev = take(d.n_nodes);
        I.o_tprev = take(d.n_nodes);
        I.o_yn = take(d.n_nodes);
        I.o_cl = take(d.num_steps);
        I.o_cdi = take(d.num_steps);
        I.o_tw = take(d.num_steps);
        I.o_tt = take(d.num_steps);
        I.o_lkey = take((int64_t)I.P * I.P);
        I.o_lfac = take((int64_t)I.P * I.P);
        I.o_lord = take(I.P);
        I.o_lst = take(1);
        sw->host.push_back(I);
    }
    // host image of the state block, then one upload
    std::vector<double> img;
    try {
        img.assign((size_t)off, 0.0);
    } catch (...) {
        delete sw;
        return set_error(ctx, FV_E_HIP, "out of host memory (%lld doubles)", (long long)off);
    }
    for (int i = 0; i < n; ++i) {
        const fv_sweep_desc &d = descs[i];
        const SwInst &I = sw->host[i];
        memcpy(&img[I.o_grid0], d.grid, sizeof(double) * 3 * I.nv);
        memcpy(&img[I.o_yn], d.y_nodes, sizeof(double) * d.n_nodes);
        const int nf = I.nfree;
        if (d.couple) {
            memcpy(&img[I.o_M], d.M, sizeof(double) * nf * nf);
            memcpy(&img[I.o_C], d.C, sizeof(double) * nf * nf);
            memcpy(&img[I.o_K], d.Keff_inv, sizeof(double) * nf * nf);
            memcpy(&img[I.o_d], d.d0, sizeof(double) * nf);
            memcpy(&img[I.o_v], d.v0, sizeof(double) * nf);
            memcpy(&img[I.o_a], d.a0, sizeof(double) * nf);
        }
        for (int64_t k = 0; k < I.num_steps; ++k) {
            img[I.o_tw + k] = NAN;
            img[I.o_tt + k] = NAN;
        }
    }
    hipError_t e1 = hipMalloc(&sw->block, sizeof(double) * (size_t)(off > 0 ? off : 1));
    hipError_t e2 = hipMalloc(&sw->dinst, sizeof(SwInst) * n);
    if (e1 != hipSuccess || e2 != hipSuccess) {
        if (sw->block) (void)hipFree(sw->block);
        if (sw->dinst) (void)hipFree(sw->dinst);
        delete sw;
        return set_error(ctx, FV_E_HIP, "hipMalloc failed for the sweep state");
    }
    int rc = FV_OK;
    if (hipMemcpyAsync(sw->block, img.data(), sizeof(double) * (size_t)off, hipMemcpyHostToDevice,
                       ctx->stream) != hipSuccess ||
        hipMemcpyAsync(sw->dinst, sw->host.data(), sizeof(SwInst) * n, hipMemcpyHostToDevice,
                       ctx->stream) != hipSuccess)
        rc = set_error(ctx, FV_E_HIP, "sweep upload failed");
    if (!rc) {
        hipLaunchKernelGGL(k_sw_init, dim3((unsigned)n), dim3(64), 0, ctx->stream, sw->dinst,
                           sw->block);
        if (hipGetLastError() != hipSuccess) rc = set_error(ctx, FV_E_HIP, "sweep init launch failed");
    }
    if (!rc && hipStreamSynchronize(ctx->stream) != hipSuccess)
        rc = set_error(ctx, FV_E_HIP, "sweep set-up failed");
    if (rc) {
        fv_sweep_destroy(sw);
        return rc;
    }
    *out = sw;
    return FV_OK;
}

void fv_sweep_destroy(fv_sweep *sw)
{
    if (!sw) return;
    (void)hipStreamSynchronize(sw->ctx->stream);
    if (sw->exec) (void)hipGraphExecDestroy(sw->exec);
    if (sw->graph) (void)hipGraphDestroy(sw->graph);
    if (sw->block) (void)hipFree(sw->block);
    if (sw->dinst) (void)hipFree(sw->dinst);
    delete sw;
}

int fv_sweep_step(fv_sweep *sw, int64_t steps)
{
    if (!sw) return FV_E_ARG;
    fv_ctx *ctx = sw->ctx;
    FV_REQUIRE(ctx, steps >= 0, "negative step count");
    int64_t k = steps;
    if (sw->done_steps + k > sw->max_steps) k = sw->max_steps - sw->done_steps;
    if (k <= 0) return FV_OK;
    if (sw->use_graph && !sw->exec) {
        FV_HIP_CHECK(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeRelaxed));
        const int rc = sweep_launch(sw);
        hipGraph_t gr = nullptr;
        const hipError_t ec = hipStreamEndCapture(ctx->stream, &gr);
        if (rc) {
            if (gr) (void)hipGraphDestroy(gr);
            return rc;
        }
        if (ec != hipSuccess) return set_error(ctx, FV_E_HIP, "sweep graph capture failed");
        sw->graph = gr;
        FV_HIP_CHECK(ctx, hipGraphInstantiate(&sw->exec, sw->graph, nullptr, nullptr, 0));
    }
    TimedLaunch tl(ctx, "sweep_step");
    for (int64_t q = 0; q < k; ++q) {
        if (sw->use_graph) {
            FV_HIP_CHECK(ctx, hipGraphLaunch(sw->exec, ctx->stream));
        } else {
            const int rc = sweep_launch(sw);
            if (rc) return rc;
        }
    }
    tl.done();
    sw->done_steps += k;
    return FV_OK;
}

int fv_sweep_read(fv_sweep *sw, int32_t i, double *cl, double *cdi, double *tip_w,
                  double *tip_theta, double *gamma, int64_t *np, int32_t *status)
{
    if (!sw) return FV_E_ARG;
    fv_ctx *ctx = sw->ctx;
    FV_REQUIRE(ctx, i >= 0 && i < sw->n, "instance %d outside 0..%d", i, sw->n - 1);
    const SwInst &H = sw->host[i];
    SwInst D;
    FV_HIP_CHECK(ctx, hipMemcpyAsync(&D, sw->dinst + i, sizeof D, hipMemcpyDeviceToHost, ctx->stream));
    const int64_t ns = H.num_steps;
    if (cl) FV_HIP_CHECK(ctx, hipMemcpyAsync(cl, sw->block + H.o_cl, sizeof(double) * ns, hipMemcpyDeviceToHost, ctx->stream));
    if (cdi) FV_HIP_CHECK(ctx, hipMemcpyAsync(cdi, sw->block + H.o_cdi, sizeof(double) * ns, hipMemcpyDeviceToHost, ctx->stream));
    if (tip_w) FV_HIP_CHECK(ctx, hipMemcpyAsync(tip_w, sw->block + H.o_tw, sizeof(double) * ns, hipMemcpyDeviceToHost, ctx->stream));
    if (tip_theta) FV_HIP_CHECK(ctx, hipMemcpyAsync(tip_theta, sw->block + H.o_tt, sizeof(double) * ns, hipMemcpyDeviceToHost, ctx->stream));
    if (gamma) FV_HIP_CHECK(ctx, hipMemcpyAsync(gamma, sw->block + H.o_gam, sizeof(double) * H.P, hipMemcpyDeviceToHost, ctx->stream));
    FV_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    if (np) *np = D.np;
    if (status) *status = D.status;
    return FV_OK;
}

}  // extern "C"
