/*
 * fv_oracle_uvlm.c -- CPU FP64 restatement of the hybrid panel-particle step
 * (TEST INFRASTRUCTURE ONLY; see fv_oracle.c for who may call it).
 *
 * Follows tests/experiment_hybrid_panel_particle.py:87-204 (HybridSolver) over
 * a ring-vortex UVLM step with PteraSoftware's data conventions
 * (idea-stage/architecture_analysis.md:5-23).  PteraSoftware itself is not
 * vendored, so the step semantics below are this project's statement of it,
 * identical to the device driver's:
 *   - rings: front legs at each panel's quarter chord, back legs at the next
 *     panel's quarter chord (last row: one panel length behind the front);
 *     collocation points at 3/4 chord, mid span; unit normal from the panel
 *     diagonals (Fr - Bl) x (Fl - Br);
 *   - motion velocity = (x(s) - x(s-1)) / dt, zero at the first step;
 *   - rhs_i = -[(v_wake.n + v_particles.n) + (V - v_motion).n];
 *   - loads: Kutta-Joukowski on the right, front and left legs with
 *     effective strengths (own minus neighbour), plus -rho A dGamma/dt n
 *     (zero at the first step); particles are not included in the loads
 *     (HybridSolver overrides only the wake-wing influence, :114-123);
 *   - wake: row 0 = trailing-edge ring back vertices of the next geometry,
 *     older rows advected by V dt, strengths shift by one row, ages + dt;
 *   - conversion of every row >= n_keep with |Gamma| >= 1e-15 (:173-200),
 *     then advect_rk3 + relaxation of all particles (:136-143).
 * Every wake row is kept (converted rows hold zero strength); rings whose
 * strength is exactly 0.0 are skipped in the sums, which adds exactly nothing.
 *
 * Pinning: the pieces are the kernel-level functions of fv_oracle.c (pinned
 * there); the step composition itself is checked by tests/test_oracle_uvlm.py
 * against PteraSoftware's published static-wing CL/CDi (reference
 * README.md:44), Helmbold's lift slope of a rectangular flat plate, linearity
 * in alpha, and the Wagner-type rise of an impulsively started plate -- the
 * CL/CDi histories of the reference's own solver are parity unpinned
 * (PteraSoftware is not available).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "fv_oracle.h"

typedef struct orc_hybrid_desc {
    int32_t n_surfaces;
    const int32_t *nc;
    const int32_t *ns;
    int64_t num_steps;
    const double *grids;
    double vinf[3];
    double rho;
    double dt;
    double nu;
    double rc0_bound;
    double rc0_wake;
    int32_t n_keep;
    double sigma;
    int32_t kernel;
    int32_t free_wake;
    int32_t stretch;
    double rlxf;
    double vpm_nu;
    double sref;
    int64_t max_particles;
    int32_t precision;   /* ignored: the oracle is FP64 */
    /* aeroelastic coupling (optional; one surface).  With `beam` set, `grids`
     * holds only the undeformed lattice: each step's lattice is that grid
     * deformed by the beam's nodal heave / twist (aeroelastic_solver.py:63-221,
     * beam_fe.py:231-351, the device sweep's statement). */
    const struct orc_beam *beam;
    /* 1: particles are a prescribed wake moved by V dt per step (no RK3 /
     * clamp / relaxation) */
    int32_t convect_exact;
    /* 1: free ring wake -- every live wake vertex moves by (V + u) dt, u the
     * bound (this step's circulation) + wake-ring induction at its pre-step
     * position, all from one snapshot (PteraSoftware's prescribed_wake=False;
     * warp_fsi/kernels_wake.py:34-108); rings whose corner-mean x exceeds
     * truncate_x lose their circulation after shedding (:110-126) */
    int32_t free_ring_wake;
    double truncate_x;
    /* particles closer than 1e-12: 0 skipped (warp_vpm.py:37,103), 1 r clamped
     * to 1e-12 as the reference's CPU field does (kernel.py:105,156) */
    int32_t pair_rule;
} orc_hybrid_desc;

typedef struct orc_beam {
    int32_t n_nodes;                /* including the clamped root */
    const double *y_nodes;
    const double *M, *C, *Keff_inv; /* free DOFs, 3 (n_nodes - 1) square, row-major */
    double *d, *v, *a;              /* free-DOF state, advanced in place */
    double beta, gamma;             /* Newmark */
    double x_ea;                    /* elastic axis x */
    double relax;                   /* deformation relaxation (1: none) */
    double *tip_w, *tip_theta;      /* [num_steps] outputs, NaN where no coupling */
    /* 1: the deformed lattice carries the aerodynamics of the next step;
     * 0: the aerodynamics stay on the undeformed lattice while the beam is
     * still driven by the loads (the reference's vertex mutation does not
     * reach its solver: its stored Goland histories show loads independent
     * of the deformation, figures/flutter_data.npz) */
    int32_t feedback;
    /* station lift = lift_sign * F_z (the reference: -F_z,
     * aeroelastic_solver.py:135-139) */
    double lift_sign;
    /* optional [num_steps x ns] outputs: station lift and moment about x = 0
     * at each coupled step (NaN elsewhere) */
    double *station_lift, *station_moment;
} orc_beam;

typedef struct {
    int nc, ns;
    int64_t poff, voff;
    /* wake: rows x (ns+1) vertices, rows x ns rings (capacity num_steps rows) */
    double *wv, *wg, *wa;
    int64_t nrow;
} surf_t;

static void lerp(const double *a, const double *b, double t, double *o)
{
    for (int c = 0; c < 3; ++c) o[c] = a[c] + t * (b[c] - a[c]);
}

static const double *gv(const double *g, int ns, int i, int j)
{
    return g + 3 * ((int64_t)i * (ns + 1) + j);
}

static void ring_of(const double *g, int nc, int ns, int i, int j, double *Br, double *Fr,
                    double *Fl, double *Bl)
{
    lerp(gv(g, ns, i, j), gv(g, ns, i + 1, j), 0.25, Fl);
    lerp(gv(g, ns, i, j + 1), gv(g, ns, i + 1, j + 1), 0.25, Fr);
    if (i < nc - 1) {
        lerp(gv(g, ns, i + 1, j), gv(g, ns, i + 2, j), 0.25, Bl);
        lerp(gv(g, ns, i + 1, j + 1), gv(g, ns, i + 2, j + 1), 0.25, Br);
    } else {
        for (int c = 0; c < 3; ++c) {
            Bl[c] = Fl[c] + (gv(g, ns, i + 1, j)[c] - gv(g, ns, i, j)[c]);
            Br[c] = Fr[c] + (gv(g, ns, i + 1, j + 1)[c] - gv(g, ns, i, j + 1)[c]);
        }
    }
}

static void frame_of(const double *g, int ns, int i, int j, double *col, double *n, double *area)
{
    double a[3], b[3];
    lerp(gv(g, ns, i, j), gv(g, ns, i + 1, j), 0.75, a);
    lerp(gv(g, ns, i, j + 1), gv(g, ns, i + 1, j + 1), 0.75, b);
    for (int c = 0; c < 3; ++c) col[c] = 0.5 * (a[c] + b[c]);
    const double *fl = gv(g, ns, i, j), *fr = gv(g, ns, i, j + 1);
    const double *bl = gv(g, ns, i + 1, j), *br = gv(g, ns, i + 1, j + 1);
    double d1[3] = {fr[0] - bl[0], fr[1] - bl[1], fr[2] - bl[2]};
    double d2[3] = {fl[0] - br[0], fl[1] - br[1], fl[2] - br[2]};
    double cr[3] = {d1[1] * d2[2] - d1[2] * d2[1], d1[2] * d2[0] - d1[0] * d2[2],
                    d1[0] * d2[1] - d1[1] * d2[0]};
    double len = sqrt(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2]);
    *area = 0.5 * len;
    for (int c = 0; c < 3; ++c) n[c] = cr[c] / len;
}

static void te_of(const double *g, int nc, int ns, int j, double *o)
{
    const double *a = gv(g, ns, nc - 1, j), *b = gv(g, ns, nc, j);
    for (int c = 0; c < 3; ++c) o[c] = b[c] + 0.25 * (b[c] - a[c]);
}

/* numpy.interp: linear inside [xp0, xp_last], the end values outside. */
static double np_interp(double x, const double *xp, const double *fp, int n)
{
    if (x <= xp[0]) return fp[0];
    if (x >= xp[n - 1]) return fp[n - 1];
    int j = 0;
    while (j + 1 < n - 1 && xp[j + 1] <= x) ++j;
    if (x == xp[j]) return fp[j];
    double slope = (fp[j + 1] - fp[j]) / (xp[j + 1] - xp[j]);
    return slope * (x - xp[j]) + fp[j];
}

/* aeroelastic_solver.py:176-198: heave w and twist theta at the vertex's |y|,
 * rotation about x_ea (applied to every lattice vertex). */
static void deform_grid(const double *base, int64_t nv, const orc_beam *bm, const double *w,
                        const double *th, double *out)
{
    for (int64_t v = 0; v < nv; ++v) {
        const double *b = base + 3 * v;
        double y = fabs(b[1]);
        double wv = np_interp(y, bm->y_nodes, w, bm->n_nodes);
        double tv = np_interp(y, bm->y_nodes, th, bm->n_nodes);
        double xr = b[0] - bm->x_ea;
        out[3 * v] = b[0] + xr * (cos(tv) - 1.0);
        out[3 * v + 1] = b[1];
        out[3 * v + 2] = b[2] + (wv + xr * sin(tv));
    }
}

/* Collects the non-zero wake rings of all surfaces into flat arrays. */
static int64_t gather_wake(surf_t *sf, int S, double *br, double *fr, double *fl, double *bl,
                           double *g, double *age)
{
    int64_t k = 0;
    for (int s = 0; s < S; ++s) {
        int ns = sf[s].ns;
        for (int64_t r = 0; r < sf[s].nrow; ++r)
            for (int j = 0; j < ns; ++j) {
                double G = sf[s].wg[r * ns + j];
                if (G == 0.0) continue;
                memcpy(fl + 3 * k, sf[s].wv + 3 * (r * (ns + 1) + j), 3 * sizeof(double));
                memcpy(fr + 3 * k, sf[s].wv + 3 * (r * (ns + 1) + j + 1), 3 * sizeof(double));
                memcpy(bl + 3 * k, sf[s].wv + 3 * ((r + 1) * (ns + 1) + j), 3 * sizeof(double));
                memcpy(br + 3 * k, sf[s].wv + 3 * ((r + 1) * (ns + 1) + j + 1),
                       3 * sizeof(double));
                g[k] = G;
                age[k] = sf[s].wa[r * ns + j];
                ++k;
            }
    }
    return k;
}

/* Runs `steps` steps from the initial state.  Outputs (any may be NULL):
 * cl/cdi [num_steps], gamma [P], particle count and state (capacity
 * max_particles).  Returns 0, -1 on allocation failure, -2 when the particle
 * field would overflow, k+1 for a singular AIC at step k. */
int orc_hybrid_run(const orc_hybrid_desc *d, int64_t steps, double *cl, double *cdi,
                   double *gamma_out, int64_t *np_out, double *pos_out, double *alpha_out,
                   double *sigma_out, int nthreads)
{
    const int S = d->n_surfaces;
    surf_t *sf = calloc((size_t)S, sizeof(surf_t));
    if (!sf) return -1;
    int64_t P = 0, TV = 0, NSt = 0;
    for (int s = 0; s < S; ++s) {
        sf[s].nc = d->nc[s];
        sf[s].ns = d->ns[s];
        sf[s].poff = P;
        sf[s].voff = TV;
        P += (int64_t)d->nc[s] * d->ns[s];
        TV += (int64_t)(d->nc[s] + 1) * (d->ns[s] + 1);
        NSt += d->ns[s];
    }
    const int64_t NS1 = d->num_steps + 1;
    int fail = 0;
    for (int s = 0; s < S; ++s) {
        sf[s].wv = calloc((size_t)(NS1 + 1) * (sf[s].ns + 1) * 3, sizeof(double));
        sf[s].wg = calloc((size_t)NS1 * sf[s].ns, sizeof(double));
        sf[s].wa = calloc((size_t)NS1 * sf[s].ns, sizeof(double));
        if (!sf[s].wv || !sf[s].wg || !sf[s].wa) fail = 1;
    }
    const int64_t WRmax = NS1 * NSt;
    double *br = malloc(sizeof(double) * 3 * P), *fr = malloc(sizeof(double) * 3 * P);
    double *fl = malloc(sizeof(double) * 3 * P), *bl = malloc(sizeof(double) * 3 * P);
    double *col = malloc(sizeof(double) * 3 * P), *nrm = malloc(sizeof(double) * 3 * P);
    double *area = malloc(sizeof(double) * P), *vmc = malloc(sizeof(double) * 3 * P);
    double *lc = malloc(sizeof(double) * 12 * P), *lv = malloc(sizeof(double) * 12 * P);
    double *vml = malloc(sizeof(double) * 12 * P), *lf = malloc(sizeof(double) * 12 * P);
    double *A = malloc(sizeof(double) * P * P), *rhs = malloc(sizeof(double) * P);
    double *gam = calloc((size_t)P, sizeof(double)), *gprev = calloc((size_t)P, sizeof(double));
    double *rc0b = malloc(sizeof(double) * P), *vw = malloc(sizeof(double) * 3 * P);
    double *vp = calloc((size_t)3 * P, sizeof(double)), *vl = malloc(sizeof(double) * 12 * P);
    double *vl2 = malloc(sizeof(double) * 12 * P);
    double *wbr = malloc(sizeof(double) * 3 * WRmax), *wfr = malloc(sizeof(double) * 3 * WRmax);
    double *wfl = malloc(sizeof(double) * 3 * WRmax), *wbl = malloc(sizeof(double) * 3 * WRmax);
    double *wg = malloc(sizeof(double) * WRmax), *wa = malloc(sizeof(double) * WRmax);
    double *wrc = malloc(sizeof(double) * WRmax), *zage = calloc((size_t)P, sizeof(double));
    const int64_t cap = d->max_particles;
    double *ppos = malloc(sizeof(double) * 3 * cap), *pal = malloc(sizeof(double) * 3 * cap);
    double *psg = malloc(sizeof(double) * cap), *page = calloc((size_t)cap, sizeof(double));
    double *tpos = malloc(sizeof(double) * 12 * NSt), *tal = malloc(sizeof(double) * 12 * NSt);
    double *tsg = malloc(sizeof(double) * 4 * NSt);
    /* aeroelastic buffers */
    const orc_beam *bm = d->beam;
    const int NN = bm ? bm->n_nodes : 1, NF = 3 * (NN - 1);
    double *pf = malloc(sizeof(double) * 3 * P);
    double *gb[3] = {NULL, NULL, NULL};   /* current, previous, next deformed lattice */
    double *wc = calloc((size_t)NN, sizeof(double)), *tc = calloc((size_t)NN, sizeof(double));
    double *wn = calloc((size_t)NN, sizeof(double)), *tn = calloc((size_t)NN, sizeof(double));
    double *wpv = calloc((size_t)NN, sizeof(double)), *tpv = calloc((size_t)NN, sizeof(double));
    double *Fn = calloc((size_t)3 * NN, sizeof(double));
    double *Fe = calloc((size_t)(NF > 0 ? NF : 1), sizeof(double));
    double *dn = calloc((size_t)(NF > 0 ? NF : 1), sizeof(double));
    double *an = calloc((size_t)(NF > 0 ? NF : 1), sizeof(double));
    int has_wprev = 0;
    if (bm) {
        for (int k = 0; k < 3; ++k) {
            gb[k] = malloc(sizeof(double) * 3 * TV);
            if (!gb[k]) fail = 1;
        }
        if (S != 1) fail = 1;
        for (int64_t k = 0; k < d->num_steps; ++k) {
            bm->tip_w[k] = NAN;
            bm->tip_theta[k] = NAN;
        }
        for (int64_t k = 0; k < d->num_steps * (int64_t)d->ns[0]; ++k) {
            if (bm->station_lift) bm->station_lift[k] = NAN;
            if (bm->station_moment) bm->station_moment[k] = NAN;
        }
    }
    if (!pf || !wc || !tc || !wn || !tn || !wpv || !tpv || !Fn || !Fe || !dn || !an) fail = 1;
    if (!br || !fr || !fl || !bl || !col || !nrm || !area || !vmc || !lc || !lv || !vml || !lf || !A ||
        !rhs || !gam || !gprev || !rc0b || !vw || !vp || !vl || !vl2 || !wbr || !wfr || !wfl ||
        !wbl || !wg || !wa || !wrc || !zage || !ppos || !pal || !psg || !page || !tpos || !tal ||
        !tsg)
        fail = 1;
    int rc = 0;
    int64_t np = 0;
    if (fail) {
        rc = -1;
        goto done;
    }
    for (int64_t i = 0; i < P; ++i) rc0b[i] = d->rc0_bound;
    for (int64_t i = 0; i < WRmax; ++i) wrc[i] = d->rc0_wake;
    const double *V = d->vinf;
    const double dt = d->dt;
    for (int64_t st = 0; st < steps && st < d->num_steps; ++st) {
        const double *G, *Gp;
        if (bm) {
            if (bm->feedback)
                deform_grid(d->grids, TV, bm, wc, tc, gb[0]);
            else
                memcpy(gb[0], d->grids, sizeof(double) * 3 * TV);
            G = gb[0];
            Gp = st > 0 ? gb[1] : NULL;
        } else {
            G = d->grids + (size_t)st * TV * 3;
            Gp = st > 0 ? d->grids + (size_t)(st - 1) * TV * 3 : NULL;
        }
        /* geometry */
        for (int s = 0; s < S; ++s) {
            const double *g = G + 3 * sf[s].voff, *gp = Gp ? Gp + 3 * sf[s].voff : NULL;
            int nc = sf[s].nc, ns = sf[s].ns;
            for (int i = 0; i < nc; ++i)
                for (int j = 0; j < ns; ++j) {
                    int64_t p = sf[s].poff + (int64_t)i * ns + j;
                    double R[4][3], Rp[4][3], c[3], cp[3], nn[3], a;
                    ring_of(g, nc, ns, i, j, R[0], R[1], R[2], R[3]);
                    frame_of(g, ns, i, j, c, nn, &a);
                    memcpy(br + 3 * p, R[0], sizeof R[0]);
                    memcpy(fr + 3 * p, R[1], sizeof R[1]);
                    memcpy(fl + 3 * p, R[2], sizeof R[2]);
                    memcpy(bl + 3 * p, R[3], sizeof R[3]);
                    memcpy(col + 3 * p, c, sizeof c);
                    memcpy(nrm + 3 * p, nn, sizeof nn);
                    area[p] = a;
                    if (gp) {
                        double np_[3], ap;
                        ring_of(gp, nc, ns, i, j, Rp[0], Rp[1], Rp[2], Rp[3]);
                        frame_of(gp, ns, i, j, cp, np_, &ap);
                    }
                    for (int k = 0; k < 3; ++k) vmc[3 * p + k] = gp ? (c[k] - cp[k]) / dt : 0.0;
                    /* legs: 0 right Br->Fr, 1 front Fr->Fl, 2 left Fl->Bl, 3 back Bl->Br */
                    for (int l = 0; l < 4; ++l) {
                        const double *s0 = R[l], *e0 = R[(l + 1) & 3];
                        for (int k = 0; k < 3; ++k) {
                            double cc = 0.5 * (s0[k] + e0[k]);
                            lc[3 * (4 * p + l) + k] = cc;
                            lv[3 * (4 * p + l) + k] = e0[k] - s0[k];
                            if (gp) {
                                double pc = 0.5 * (Rp[l][k] + Rp[(l + 1) & 3][k]);
                                vml[3 * (4 * p + l) + k] = (cc - pc) / dt;
                            } else {
                                vml[3 * (4 * p + l) + k] = 0.0;
                            }
                        }
                    }
                }
        }
        /* AIC */
        orc_aic(col, nrm, P, br, fr, fl, bl, rc0b, d->nu, A, nthreads);
        /* right-hand side */
        int64_t W = gather_wake(sf, S, wbr, wfr, wfl, wbl, wg, wa);
        orc_collapsed(col, P, wbr, wfr, wfl, wbl, W, wg, wrc, wa, d->nu, 0xF, vw, nthreads);
        if (np > 0)
            orc_particle_vel(col, P, ppos, pal, psg, np, d->kernel, vp, nthreads);
        for (int64_t i = 0; i < P; ++i) {
            const double *n = nrm + 3 * i;
            double fs = (V[0] - vmc[3 * i]) * n[0] + (V[1] - vmc[3 * i + 1]) * n[1] +
                        (V[2] - vmc[3 * i + 2]) * n[2];
            double w = vw[3 * i] * n[0] + vw[3 * i + 1] * n[1] + vw[3 * i + 2] * n[2];
            double q = np > 0 ? vp[3 * i] * n[0] + vp[3 * i + 1] * n[1] + vp[3 * i + 2] * n[2]
                              : 0.0;
            rhs[i] = -((w + q) + fs);
        }
        int info = orc_lu_solve(P, A, rhs);
        if (info) {
            rc = (int)(st + 1);
            goto done;
        }
        memcpy(gam, rhs, sizeof(double) * P);
        /* loads */
        orc_collapsed(lc, 4 * P, br, fr, fl, bl, P, gam, rc0b, zage, d->nu, 0xF, vl, nthreads);
        orc_collapsed(lc, 4 * P, wbr, wfr, wfl, wbl, W, wg, wrc, wa, d->nu, 0xF, vl2, nthreads);
        /* Kutta-Joukowski on each ring's legs with the net circulation of the
         * line it lies on: right / left legs against the spanwise neighbour
         * (each side counts its own), the front leg against the ring ahead
         * (whole on the leading edge), the back leg only on the trailing edge,
         * against the newest wake ring (whole when there is none) -- interior
         * back legs are the next ring's front legs.  Plus the unsteady term
         * -rho A dGamma/dt n. */
        double F[3] = {0.0, 0.0, 0.0};
        for (int s = 0; s < S; ++s) {
            int ns = sf[s].ns, nc = sf[s].nc;
            for (int64_t loc = 0; loc < (int64_t)nc * ns; ++loc) {
                int64_t p = sf[s].poff + loc;
                int i = (int)(loc / ns), j = (int)(loc % ns);
                double Gm = gam[p];
                double gw = sf[s].nrow > 0 ? sf[s].wg[j] : 0.0;
                double ge[4] = {(j + 1 < ns) ? Gm - gam[p + 1] : Gm, (i > 0) ? Gm - gam[p - ns] : Gm,
                                (j > 0) ? Gm - gam[p - 1] : Gm, (i == nc - 1) ? Gm - gw : 0.0};
                double Fp[3] = {0.0, 0.0, 0.0};
                for (int l = 0; l < 4; ++l) {
                    int64_t q = 4 * p + l;
                    double u[3];
                    for (int k = 0; k < 3; ++k)
                        u[k] = V[k] - vml[3 * q + k] + (vl[3 * q + k] + vl2[3 * q + k]);
                    const double *L = lv + 3 * q;
                    double f = d->rho * ge[l];
                    lf[3 * q] = f * (u[1] * L[2] - u[2] * L[1]);
                    lf[3 * q + 1] = f * (u[2] * L[0] - u[0] * L[2]);
                    lf[3 * q + 2] = f * (u[0] * L[1] - u[1] * L[0]);
                    for (int k = 0; k < 3; ++k) Fp[k] += lf[3 * q + k];
                }
                double dG = st == 0 ? 0.0 : (Gm - gprev[p]) / dt;
                double fu = -d->rho * area[p] * dG;
                for (int k = 0; k < 3; ++k) F[k] += Fp[k] + fu * nrm[3 * p + k];
                /* the panel's own force: the line between two rows of a
                 * strip is shared half and half by the two panels (fitted
                 * to the reference's stored Goland station moments) */
                for (int k = 0; k < 3; ++k) {
                    double fr_ = lf[3 * (4 * p + 1) + k] * (i > 0 ? 0.5 : 1.0);
                    pf[3 * p + k] = lf[3 * (4 * p) + k] + lf[3 * (4 * p + 2) + k] + fr_ +
                                    lf[3 * (4 * p + 3) + k] + fu * nrm[3 * p + k];
                }
            }
            for (int64_t loc = ns; loc < (int64_t)nc * ns; ++loc) {
                int64_t p = sf[s].poff + loc;
                for (int k = 0; k < 3; ++k) pf[3 * (p - ns) + k] += 0.5 * lf[3 * (4 * p + 1) + k];
            }
        }
        {
            double V2 = V[0] * V[0] + V[1] * V[1] + V[2] * V[2], Vn = sqrt(V2);
            double dd[3] = {V[0] / Vn, V[1] / Vn, V[2] / Vn};
            double l0 = -dd[2], l2 = dd[0], ln = sqrt(l0 * l0 + l2 * l2);
            l0 /= ln;
            l2 /= ln;
            double qs = 0.5 * d->rho * V2 * d->sref;
            if (cl) cl[st] = (F[0] * l0 + F[2] * l2) / qs;
            if (cdi) cdi[st] = (F[0] * dd[0] + F[1] * dd[1] + F[2] * dd[2]) / qs;
            /* structure (aeroelastic_solver.py:68-110, beam_fe.py:231-351) */
            if (bm && st >= 1 && st <= d->num_steps - 2) {
                const int nc = sf[0].nc, ns = sf[0].ns;
                for (int k = 0; k < 3 * NN; ++k) Fn[k] = 0.0;
                for (int j = 0; j < ns; ++j) {
                    double lift = 0.0, mom = 0.0;
                    for (int i = 0; i < nc; ++i) {
                        int64_t p = (int64_t)i * ns + j;
                        double lp = bm->lift_sign * pf[3 * p + 2];
                        lift += lp;
                        mom += col[3 * p] * lp;
                    }
                    double y = col[3 * j + 1];
                    if (bm->station_lift) bm->station_lift[st * ns + j] = lift;
                    if (bm->station_moment) bm->station_moment[st * ns + j] = mom;
                    if (!(fabs(y) > 1e-6)) continue;
                    int idx = -1;
                    for (int k = 0; k < NN; ++k)
                        if (bm->y_nodes[k] <= y) idx = k;
                    idx = idx < 0 ? 0 : (idx > NN - 2 ? NN - 2 : idx);
                    double Le = bm->y_nodes[idx + 1] - bm->y_nodes[idx];
                    double xi = (y - bm->y_nodes[idx]) / Le;
                    xi = xi < 0.0 ? 0.0 : (xi > 1.0 ? 1.0 : xi);
                    Fn[3 * idx] += lift * (1 - xi);
                    Fn[3 * (idx + 1)] += lift * xi;
                    Fn[3 * idx + 2] += mom * (1 - xi);
                    Fn[3 * (idx + 1) + 2] += mom * xi;
                }
                const double be = bm->beta, ga = bm->gamma;
                const double a0 = 1.0 / (be * dt * dt), a1 = 1.0 / (be * dt),
                             a2 = 1.0 / (2.0 * be) - 1.0;
                const double a3 = ga / (be * dt), a4 = ga / be - 1.0,
                             a5 = dt * (ga / (2.0 * be) - 1.0);
                double *dv = bm->d, *vv = bm->v, *av = bm->a;
                for (int r = 0; r < NF; ++r) {
                    double sm = 0.0, sc = 0.0;
                    for (int c = 0; c < NF; ++c) {
                        sm += bm->M[r * NF + c] * (a0 * dv[c] + a1 * vv[c] + a2 * av[c]);
                        sc += bm->C[r * NF + c] * (a3 * dv[c] + a4 * vv[c] + a5 * av[c]);
                    }
                    Fe[r] = Fn[3 + r] + sm + sc;
                }
                for (int r = 0; r < NF; ++r) {
                    double x = 0.0;
                    for (int c = 0; c < NF; ++c) x += bm->Keff_inv[r * NF + c] * Fe[c];
                    dn[r] = x;
                    an[r] = a0 * (dn[r] - dv[r]) - a1 * vv[r] - a2 * av[r];
                }
                for (int r = 0; r < NF; ++r) {
                    vv[r] = vv[r] + dt * ((1.0 - ga) * av[r] + ga * an[r]);
                    av[r] = an[r];
                    dv[r] = dn[r];
                }
                for (int k = 0; k < NN; ++k) {
                    double w = k == 0 ? 0.0 : dv[3 * (k - 1)];
                    double th = k == 0 ? 0.0 : dv[3 * (k - 1) + 2];
                    if (has_wprev) {
                        w = bm->relax * w + (1.0 - bm->relax) * wpv[k];
                        th = bm->relax * th + (1.0 - bm->relax) * tpv[k];
                    }
                    wn[k] = w;
                    tn[k] = th;
                    wpv[k] = w;
                    tpv[k] = th;
                }
                has_wprev = 1;
                bm->tip_w[st] = wn[NN - 1];
                bm->tip_theta[st] = tn[NN - 1];
            } else if (bm) {
                for (int k = 0; k < NN; ++k) {
                    wn[k] = 0.0;
                    tn[k] = 0.0;
                }
            }
        }
        memcpy(gprev, gam, sizeof(double) * P);
        /* wake shedding + conversion */
        if (st + 1 < d->num_steps) {
            const double *Gn;
            if (bm) {
                if (bm->feedback)
                    deform_grid(d->grids, TV, bm, wn, tn, gb[2]);
                else
                    memcpy(gb[2], d->grids, sizeof(double) * 3 * TV);
                Gn = gb[2];
            } else {
                Gn = d->grids + (size_t)(st + 1) * TV * 3;
            }
            int64_t added = 0;
            for (int s = 0; s < S; ++s) {
                int nc = sf[s].nc, ns = sf[s].ns;
                int64_t nr = sf[s].nrow;
                /* free ring wake: induced velocity at the live old vertex rows
                 * 0 .. min(nr, n_keep) (rows beyond hold no circulation) */
                const int64_t nfree = d->free_ring_wake ? (nr < d->n_keep ? nr : d->n_keep) + 1 : 0;
                double *fp = NULL, *fu = NULL, *fu2 = NULL;
                if (nfree > 0) {
                    const int64_t npt = nfree * (ns + 1);
                    fp = malloc(sizeof(double) * 3 * npt);
                    fu = malloc(sizeof(double) * 3 * npt);
                    fu2 = malloc(sizeof(double) * 3 * npt);
                    if (!fp || !fu || !fu2) {
                        free(fp); free(fu); free(fu2);
                        rc = -1;
                        goto done;
                    }
                    for (int64_t r = 0; r < nfree; ++r)
                        for (int j = 0; j <= ns; ++j) {
                            double *q = fp + 3 * (r * (ns + 1) + j);
                            if (r == 0)
                                te_of(G + 3 * sf[s].voff, nc, ns, j, q);
                            else
                                memcpy(q, sf[s].wv + 3 * (r * (ns + 1) + j), 3 * sizeof(double));
                        }
                    orc_collapsed(fp, npt, br, fr, fl, bl, P, gam, rc0b, zage, d->nu, 0xF, fu,
                                  nthreads);
                    orc_collapsed(fp, npt, wbr, wfr, wfl, wbl, W, wg, wrc, wa, d->nu, 0xF, fu2,
                                  nthreads);
                }
                /* shift rows back-to-front in place */
                for (int64_t r = nr; r >= 0; --r)
                    for (int j = 0; j <= ns; ++j) {
                        double v[3];
                        if (r == 0)
                            te_of(G + 3 * sf[s].voff, nc, ns, j, v);
                        else
                            memcpy(v, sf[s].wv + 3 * (r * (ns + 1) + j), sizeof v);
                        /* old vertex row r (r = 0: current trailing edge) -> row r+1 */
                        double *o = sf[s].wv + 3 * ((r + 1) * (ns + 1) + j);
                        if (r < nfree) {
                            const int64_t q = 3 * (r * (ns + 1) + j);
                            for (int k = 0; k < 3; ++k)
                                o[k] = v[k] + (V[k] + (fu[q + k] + fu2[q + k])) * dt;
                        } else {
                            for (int k = 0; k < 3; ++k) o[k] = v[k] + V[k] * dt;
                        }
                    }
                free(fp); free(fu); free(fu2);
                for (int j = 0; j <= ns; ++j) te_of(Gn + 3 * sf[s].voff, nc, ns, j,
                                                    sf[s].wv + 3 * j);
                for (int64_t r = nr; r >= 1; --r)
                    for (int j = 0; j < ns; ++j) {
                        sf[s].wg[r * ns + j] = sf[s].wg[(r - 1) * ns + j];
                        sf[s].wa[r * ns + j] = sf[s].wa[(r - 1) * ns + j] + dt;
                    }
                for (int j = 0; j < ns; ++j) {
                    sf[s].wg[j] = gam[sf[s].poff + (int64_t)(nc - 1) * ns + j];
                    sf[s].wa[j] = 0.0;
                }
                sf[s].nrow = nr + 1;
            }
            /* truncation by ring corner-mean x (Fr + Fl + Bl + Br) / 4 */
            if (isfinite(d->truncate_x))
                for (int s = 0; s < S; ++s) {
                    const int ns = sf[s].ns;
                    for (int64_t r = 0; r < sf[s].nrow; ++r)
                        for (int j = 0; j < ns; ++j) {
                            const double *a_ = sf[s].wv + 3 * (r * (ns + 1) + j), *b_ = a_ + 3;
                            const double *c_ = sf[s].wv + 3 * ((r + 1) * (ns + 1) + j), *e_ = c_ + 3;
                            const double cx = (b_[0] + a_[0] + c_[0] + e_[0]) * 0.25;
                            if (cx > d->truncate_x) sf[s].wg[r * ns + j] = 0.0;
                        }
                }
            /* conversion of rows >= n_keep (surface, row, span order) */
            for (int s = 0; s < S; ++s) {
                int ns = sf[s].ns;
                for (int64_t r = d->n_keep; r < sf[s].nrow; ++r)
                    for (int j = 0; j < ns; ++j) {
                        double *gg = sf[s].wg + r * ns + j;
                        if (fabs(*gg) < 1e-15) continue;
                        double Fl_[3], Fr_[3], Bl_[3], Br_[3];
                        memcpy(Fl_, sf[s].wv + 3 * (r * (ns + 1) + j), sizeof Fl_);
                        memcpy(Fr_, sf[s].wv + 3 * (r * (ns + 1) + j + 1), sizeof Fr_);
                        memcpy(Bl_, sf[s].wv + 3 * ((r + 1) * (ns + 1) + j), sizeof Bl_);
                        memcpy(Br_, sf[s].wv + 3 * ((r + 1) * (ns + 1) + j + 1), sizeof Br_);
                        double pp[12], aa[12], ss[4];
                        int64_t k = orc_convert_rings(1, Br_, Fr_, Fl_, Bl_, gg, d->sigma, pp, aa,
                                                      ss);
                        if (np + added + k > cap) {
                            rc = -2;
                            goto done;
                        }
                        memcpy(ppos + 3 * (np + added), pp, sizeof(double) * 3 * k);
                        memcpy(pal + 3 * (np + added), aa, sizeof(double) * 3 * k);
                        for (int64_t q = 0; q < k; ++q) {
                            psg[np + added + q] = ss[q];
                            page[np + added + q] = 0.0;
                        }
                        added += k;
                    }
            }
            np += added;
        }
        if (np > 0 && d->convect_exact) {
            for (int64_t k = 0; k < np; ++k)
                for (int c = 0; c < 3; ++c) ppos[3 * k + c] += V[c] * dt;
        } else if (np > 0) {
            orc_vpm_advect_rk3_rule(np, ppos, pal, psg, page, dt, V, d->stretch, d->free_wake,
                                    d->kernel, d->vpm_nu, d->pair_rule, nthreads);
            orc_vpm_relax_rule(np, ppos, pal, psg, d->rlxf, d->pair_rule, nthreads);
        }
        if (bm) {
            /* this lattice becomes the previous one; the next deformation the current */
            double *tg = gb[1];
            gb[1] = gb[0];
            gb[0] = tg;
            memcpy(wc, wn, sizeof(double) * NN);
            memcpy(tc, tn, sizeof(double) * NN);
        }
    }
    if (gamma_out) memcpy(gamma_out, gam, sizeof(double) * P);
    if (np_out) *np_out = np;
    if (pos_out) memcpy(pos_out, ppos, sizeof(double) * 3 * np);
    if (alpha_out) memcpy(alpha_out, pal, sizeof(double) * 3 * np);
    if (sigma_out) memcpy(sigma_out, psg, sizeof(double) * np);
done:
    for (int s = 0; s < S; ++s) {
        free(sf[s].wv);
        free(sf[s].wg);
        free(sf[s].wa);
    }
    free(sf);
    free(br); free(fr); free(fl); free(bl); free(col); free(nrm); free(area); free(vmc);
    free(lc); free(lv); free(vml); free(lf); free(A); free(rhs); free(gam); free(gprev); free(rc0b);
    free(vw); free(vp); free(vl); free(vl2); free(wbr); free(wfr); free(wfl); free(wbl);
    free(wg); free(wa); free(wrc); free(zage); free(ppos); free(pal); free(psg); free(page);
    free(tpos); free(tal); free(tsg);
    free(pf); free(gb[0]); free(gb[1]); free(gb[2]); free(wc); free(tc); free(wn); free(tn);
    free(wpv); free(tpv); free(Fn); free(Fe); free(dn); free(an);
    return rc;
}
