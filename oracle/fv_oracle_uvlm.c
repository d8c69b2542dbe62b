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
} orc_hybrid_desc;

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
    double *lc = malloc(sizeof(double) * 9 * P), *lv = malloc(sizeof(double) * 9 * P);
    double *vml = malloc(sizeof(double) * 9 * P);
    double *A = malloc(sizeof(double) * P * P), *rhs = malloc(sizeof(double) * P);
    double *gam = calloc((size_t)P, sizeof(double)), *gprev = calloc((size_t)P, sizeof(double));
    double *rc0b = malloc(sizeof(double) * P), *vw = malloc(sizeof(double) * 3 * P);
    double *vp = calloc((size_t)3 * P, sizeof(double)), *vl = malloc(sizeof(double) * 9 * P);
    double *vl2 = malloc(sizeof(double) * 9 * P);
    double *wbr = malloc(sizeof(double) * 3 * WRmax), *wfr = malloc(sizeof(double) * 3 * WRmax);
    double *wfl = malloc(sizeof(double) * 3 * WRmax), *wbl = malloc(sizeof(double) * 3 * WRmax);
    double *wg = malloc(sizeof(double) * WRmax), *wa = malloc(sizeof(double) * WRmax);
    double *wrc = malloc(sizeof(double) * WRmax), *zage = calloc((size_t)P, sizeof(double));
    const int64_t cap = d->max_particles;
    double *ppos = malloc(sizeof(double) * 3 * cap), *pal = malloc(sizeof(double) * 3 * cap);
    double *psg = malloc(sizeof(double) * cap), *page = calloc((size_t)cap, sizeof(double));
    double *tpos = malloc(sizeof(double) * 12 * NSt), *tal = malloc(sizeof(double) * 12 * NSt);
    double *tsg = malloc(sizeof(double) * 4 * NSt);
    if (!br || !fr || !fl || !bl || !col || !nrm || !area || !vmc || !lc || !lv || !vml || !A ||
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
        const double *G = d->grids + (size_t)st * TV * 3;
        const double *Gp = st > 0 ? d->grids + (size_t)(st - 1) * TV * 3 : NULL;
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
                    /* legs: 0 right Br->Fr, 1 front Fr->Fl, 2 left Fl->Bl */
                    for (int l = 0; l < 3; ++l) {
                        const double *s0 = R[l], *e0 = R[l + 1];
                        for (int k = 0; k < 3; ++k) {
                            double cc = 0.5 * (s0[k] + e0[k]);
                            lc[3 * (3 * p + l) + k] = cc;
                            lv[3 * (3 * p + l) + k] = e0[k] - s0[k];
                            if (gp) {
                                double pc = 0.5 * (Rp[l][k] + Rp[l + 1][k]);
                                vml[3 * (3 * p + l) + k] = (cc - pc) / dt;
                            } else {
                                vml[3 * (3 * p + l) + k] = 0.0;
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
        orc_collapsed(lc, 3 * P, br, fr, fl, bl, P, gam, rc0b, zage, d->nu, 0xF, vl, nthreads);
        orc_collapsed(lc, 3 * P, wbr, wfr, wfl, wbl, W, wg, wrc, wa, d->nu, 0xF, vl2, nthreads);
        double F[3] = {0.0, 0.0, 0.0};
        for (int s = 0; s < S; ++s) {
            int ns = sf[s].ns;
            for (int64_t loc = 0; loc < (int64_t)sf[s].nc * ns; ++loc) {
                int64_t p = sf[s].poff + loc;
                int i = (int)(loc / ns), j = (int)(loc % ns);
                double Gm = gam[p];
                double ge[3] = {(j + 1 < ns) ? Gm - gam[p + 1] : Gm, (i > 0) ? Gm - gam[p - ns] : Gm,
                                (j > 0) ? Gm - gam[p - 1] : Gm};
                double Fp[3] = {0.0, 0.0, 0.0};
                for (int l = 0; l < 3; ++l) {
                    int64_t q = 3 * p + l;
                    double u[3];
                    for (int k = 0; k < 3; ++k)
                        u[k] = V[k] - vml[3 * q + k] + (vl[3 * q + k] + vl2[3 * q + k]);
                    const double *L = lv + 3 * q;
                    double f = d->rho * ge[l];
                    Fp[0] += f * (u[1] * L[2] - u[2] * L[1]);
                    Fp[1] += f * (u[2] * L[0] - u[0] * L[2]);
                    Fp[2] += f * (u[0] * L[1] - u[1] * L[0]);
                }
                double dG = st == 0 ? 0.0 : (Gm - gprev[p]) / dt;
                double fu = -d->rho * area[p] * dG;
                for (int k = 0; k < 3; ++k) F[k] += Fp[k] + fu * nrm[3 * p + k];
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
        }
        memcpy(gprev, gam, sizeof(double) * P);
        /* wake shedding + conversion */
        if (st + 1 < d->num_steps) {
            const double *Gn = d->grids + (size_t)(st + 1) * TV * 3;
            int64_t added = 0;
            for (int s = 0; s < S; ++s) {
                int nc = sf[s].nc, ns = sf[s].ns;
                int64_t nr = sf[s].nrow;
                /* shift rows back-to-front in place */
                for (int64_t r = nr; r >= 0; --r)
                    for (int j = 0; j <= ns; ++j) {
                        double v[3];
                        if (r == 0)
                            te_of(G + 3 * sf[s].voff, nc, ns, j, v);
                        else
                            memcpy(v, sf[s].wv + 3 * (r * (ns + 1) + j), sizeof v);
                        /* old vertex row r (r = 0: current trailing edge) -> row r+1 */
                        for (int k = 0; k < 3; ++k)
                            sf[s].wv[3 * ((r + 1) * (ns + 1) + j) + k] = v[k] + V[k] * dt;
                    }
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
        if (np > 0) {
            orc_vpm_advect_rk3(np, ppos, pal, psg, page, dt, V, d->stretch, d->free_wake,
                               d->kernel, d->vpm_nu, nthreads);
            orc_vpm_relax(np, ppos, pal, psg, d->rlxf, nthreads);
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
    free(lc); free(lv); free(vml); free(A); free(rhs); free(gam); free(gprev); free(rc0b);
    free(vw); free(vp); free(vl); free(vl2); free(wbr); free(wfr); free(wfl); free(wbl);
    free(wg); free(wa); free(wrc); free(zage); free(ppos); free(pal); free(psg); free(page);
    free(tpos); free(tal); free(tsg);
    return rc;
}
