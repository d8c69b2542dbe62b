/*
 * fv_oracle.c -- CPU FP64 restatement of FLUXVortex's hybrid panel-particle
 * hot path.  TEST INFRASTRUCTURE ONLY: this file is the checker.  Only the
 * tests, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it;
 * the product library (fluxvortex_amd) never links, loads or calls it.
 *
 * Built with -O3 -ffp-contract=off (no FMA contraction, so every expression is
 * evaluated in the order the reference's NumPy/Warp code writes it) and OpenMP
 * over targets (each target's sum runs in a fixed serial order, so results do
 * not depend on the thread count).
 *
 * Reference files restated (paths relative to the reference repository):
 *   src/fluxvortex/warp_kernels.py:16-382      line-vortex Biot-Savart
 *   newton_pc/adapters/flap.py:195-228         NumPy Ptera-form segment kernel
 *   src/fluxvortex/kernel.py:1-190             particle velocity / Jacobian
 *   src/fluxvortex/warp_vpm.py:19-173          self-pair rule (r < 1e-12 skipped)
 *   src/fluxvortex/particles.py:15-246         RK3 + rVPM + Pedrizzetti relaxation
 *   tests/experiment_hybrid_panel_particle.py:145-204   N_keep conversion
 *   src/fluxvortex/standalone_uvlm.py:390-437  dense circulation solve semantics
 *
 * Pinning (what checks each function; see tests/test_oracle.py):
 *   seg_vel_acc / orc_collapsed / orc_expanded : closed forms (finite segment,
 *       square ring centre, infinite-line limit), linearity, collapsed == sum of
 *       expanded, leg decomposition, on-line/zero-length zeros, div v = 0 by
 *       finite differences, and an independent NumPy restatement.
 *   orc_particle_vel / orc_particle_jac : Winckelmans g(0.5) = 0.19677, far-field
 *       point-vortex limit, Jacobian vs central differences of the velocity,
 *       trace J = 0 at sigma -> 0, NumPy restatement.
 *   orc_lu_solve : residual vs numpy.linalg.solve.
 *   orc_vpm_advect_rk3 / orc_vpm_advect_rk3_ext : constant field moves
 *       particles by exactly U dt, each clamp forced one at a time; whole
 *       trajectories written by the reference's own particles.py
 *       (tests/golden/ref_vpm.npz, tests/test_ref_golden.py: the stage
 *       composition, the clamped self pair, the per-stage callables), which a
 *       planted stage-order mistake or a freestream frozen per step fails.
 *   orc_vpm_relax : fixed points (Gamma parallel to omega), unit-norm limit,
 *       and the same reference trajectories with rlxf = 0.3.
 *   orc_convert_rings : sum of alpha over a closed ring = 0, leg midpoints,
 *       far-field dipole equivalence.
 */
#include "fv_oracle.h"
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_INV_4PI 0.07957747154594767 /* warp_kernels.py:20 */
#define ORC_FOUR_LAMB 5.02572           /* warp_kernels.py:18, 4 * 1.25643 */
#define ORC_SQUIRE 1.0e-4               /* warp_kernels.py:19 */

static int orc_threads(int nthreads)
{
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#else
    nthreads = 1;
#endif
    return nthreads;
}

/* ------------------------------------------------------------------------ */
/* Line-vortex segment, Ptera form.  warp_kernels.py:46-96                  */
/* ------------------------------------------------------------------------ */
static inline void seg_vel_acc(const double *p, const double *s, const double *e,
                               double strength, double rc0, double age, double nu,
                               double *acc)
{
    double r0x = e[0] - s[0];
    double r0y = e[1] - s[1];
    double r0z = e[2] - s[2];
    double r0_sq = r0x * r0x + r0y * r0y + r0z * r0z;
    double r0_len = sqrt(r0_sq);
    if (r0_len < 2.22e-16) return;                       /* :52-53 */

    double r1x = s[0] - p[0], r1y = s[1] - p[1], r1z = s[2] - p[2];
    double r2x = e[0] - p[0], r2y = e[1] - p[1], r2z = e[2] - p[2];
    double r3x = r1y * r2z - r1z * r2y;                  /* :63-65 */
    double r3y = r1z * r2x - r1x * r2z;
    double r3z = r1x * r2y - r1y * r2x;
    double r1_len = sqrt(r1x * r1x + r1y * r1y + r1z * r1z);
    double r2_len = sqrt(r2x * r2x + r2y * r2y + r2z * r2z);
    double r0_tol = r0_len * 1.0e-10;
    if (r1_len < r0_tol) return;                         /* :72-75 */
    if (r2_len < r0_tol) return;
    double r3_sq = r3x * r3x + r3y * r3y + r3z * r3z;
    double r3_len = sqrt(r3_sq);
    double r1r2 = r1_len * r2_len;
    if (r3_len < 1.0e-10 * r1r2) return;                 /* :81-82 */

    /* Lamb-Oseen + Squire core growth, :88 */
    double rc_sq = rc0 * rc0 + ORC_FOUR_LAMB * (nu + ORC_SQUIRE * fabs(strength)) * age;
    double c1 = strength * ORC_INV_4PI;
    double c2 = r0_sq * rc_sq;
    double r1pr2 = r1_len + r2_len;
    double c3 = r1x * r2x + r1y * r2y + r1z * r2z;
    double c4 = c1 * r1pr2 * (r1r2 - c3) / (r1r2 * (r3_sq + c2));   /* :94 */
    acc[0] += c4 * r3x;
    acc[1] += c4 * r3y;
    acc[2] += c4 * r3z;
}

/* Leg k runs from corner k to corner (k+1)%4 with corners (Br, Fr, Fl, Bl):
 * leg 0 Br->Fr, 1 Fr->Fl, 2 Fl->Bl, 3 Bl->Br  (warp_kernels.py:268-273).
 * Rings use mask 0xF, chordwise segments 0x5 (:296-299), horseshoes 0x7
 * (:350-354). */

/* Collapsed (summed over vortices and legs), warp_kernels.py:188-215.
 * out is (N,3), overwritten.  age may be NULL (zeros, :265-266). */
int orc_collapsed(const double *P, int64_t N,
                  const double *Br, const double *Fr, const double *Fl, const double *Bl,
                  int64_t M, const double *gamma, const double *rc0, const double *age,
                  double nu, int leg_mask, double *out, int nthreads)
{
    const double *c[4] = {Br, Fr, Fl, Bl};
    memset(out, 0, sizeof(double) * 3 * (size_t)N);
    if (M == 0 || N == 0) return 0;                      /* :193-194 */
#pragma omp parallel for schedule(static) num_threads(orc_threads(nthreads))
    for (int64_t i = 0; i < N; ++i) {
        double acc[3] = {0.0, 0.0, 0.0};
        for (int leg = 0; leg < 4; ++leg) {
            if (!(leg_mask & (1 << leg))) continue;
            const double *S = c[leg], *E = c[(leg + 1) & 3];
            for (int64_t m = 0; m < M; ++m)
                seg_vel_acc(P + 3 * i, S + 3 * m, E + 3 * m, gamma[m], rc0[m],
                            age ? age[m] : 0.0, nu, acc);
        }
        out[3 * i + 0] = acc[0];
        out[3 * i + 1] = acc[1];
        out[3 * i + 2] = acc[2];
    }
    return 0;
}

/* Expanded (per point, per vortex; legs summed), warp_kernels.py:219-246.
 * out is (N,M,3), overwritten. */
int orc_expanded(const double *P, int64_t N,
                 const double *Br, const double *Fr, const double *Fl, const double *Bl,
                 int64_t M, const double *gamma, const double *rc0, const double *age,
                 double nu, int leg_mask, double *out, int nthreads)
{
    const double *c[4] = {Br, Fr, Fl, Bl};
    memset(out, 0, sizeof(double) * 3 * (size_t)N * (size_t)M);
    if (M == 0 || N == 0) return 0;
#pragma omp parallel for schedule(static) num_threads(orc_threads(nthreads))
    for (int64_t i = 0; i < N; ++i) {
        for (int64_t m = 0; m < M; ++m) {
            double acc[3] = {0.0, 0.0, 0.0};
            for (int leg = 0; leg < 4; ++leg) {
                if (!(leg_mask & (1 << leg))) continue;
                seg_vel_acc(P + 3 * i, c[leg] + 3 * m, c[(leg + 1) & 3] + 3 * m, gamma[m],
                            rc0[m], age ? age[m] : 0.0, nu, acc);
            }
            double *o = out + 3 * ((size_t)i * (size_t)M + (size_t)m);
            o[0] = acc[0];
            o[1] = acc[1];
            o[2] = acc[2];
        }
    }
    return 0;
}

/* AIC[i][j] = (sum over the 4 legs of ring j, unit strength, zero age) . n_i.
 * Ptera's _calculate_wing_wing_influences = expanded(colloc, rings, ones) then
 * einsum('ijk,ik->ij') with the unit normals (idea-stage/architecture_analysis.md:12-13). */
int orc_aic(const double *colloc, const double *normals, int64_t P,
            const double *Br, const double *Fr, const double *Fl, const double *Bl,
            const double *rc0, double nu, double *aic, int nthreads)
{
    const double *c[4] = {Br, Fr, Fl, Bl};
#pragma omp parallel for schedule(static) num_threads(orc_threads(nthreads))
    for (int64_t i = 0; i < P; ++i) {
        for (int64_t j = 0; j < P; ++j) {
            double acc[3] = {0.0, 0.0, 0.0};
            for (int leg = 0; leg < 4; ++leg)
                seg_vel_acc(colloc + 3 * i, c[leg] + 3 * j, c[(leg + 1) & 3] + 3 * j, 1.0,
                            rc0[j], 0.0, nu, acc);
            const double *n = normals + 3 * i;
            aic[i * P + j] = acc[0] * n[0] + acc[1] * n[1] + acc[2] * n[2];
        }
    }
    return 0;
}

/* Dense solve A x = b by LU with partial pivoting (LAPACK dgesv semantics,
 * as np.linalg.solve in standalone_uvlm.py:427).  A is row-major n x n and is
 * overwritten by its factors; b is overwritten by x.  Returns 0, or k+1 when
 * the pivot in column k is exactly zero (singular). */
int orc_lu_solve(int64_t n, double *A, double *b)
{
    for (int64_t k = 0; k < n; ++k) {
        int64_t piv = k;
        double best = fabs(A[k * n + k]);
        for (int64_t r = k + 1; r < n; ++r) {
            double v = fabs(A[r * n + k]);
            if (v > best) { best = v; piv = r; }
        }
        if (best == 0.0) return (int)(k + 1);
        if (piv != k) {
            for (int64_t c = 0; c < n; ++c) {
                double t = A[k * n + c]; A[k * n + c] = A[piv * n + c]; A[piv * n + c] = t;
            }
            double t = b[k]; b[k] = b[piv]; b[piv] = t;
        }
        double inv = 1.0 / A[k * n + k];
        for (int64_t r = k + 1; r < n; ++r) {
            double l = A[r * n + k] * inv;
            A[r * n + k] = l;
            for (int64_t c = k + 1; c < n; ++c) A[r * n + c] -= l * A[k * n + c];
            b[r] -= l * b[k];
        }
    }
    for (int64_t k = n - 1; k >= 0; --k) {
        double s = b[k];
        for (int64_t c = k + 1; c < n; ++c) s -= A[k * n + c] * b[c];
        b[k] = s / A[k * n + k];
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Vortex particles.  kernel.py:10-190, warp_vpm.py:19-141                   */
/* kernel id: 0 = Gaussian-erf (kernel.py:26-31), 1 = Winckelmans (:51-55)  */
/* ------------------------------------------------------------------------ */
/* erf as scipy.special.erf evaluates it (SciPy 1.15.3, the library kernel.py
 * imports; its erf is Moshier's Cephes ndtr.c): for |x| < 1 the rational
 * x T(x^2) / U(x^2) with Horner evaluation (polevl / p1evl), coefficients
 * as published in Cephes.  Bit-identical to scipy there (checked over 2e5
 * arguments by tests/test_oracle.py); beyond, libm's erf (within an ulp, no
 * cancellation in g for rho >= 1.4).  The bits matter where g cancels: at the
 * clamped self pair, r = 1e-12, g is pure round-off whose sign decides the
 * direction the relaxation turns a particle. */
static double orc_erf_ref(double x)
{
    if (fabs(x) >= 1.0) return erf(x);
    static const double T[5] = {9.60497373987051638749E0, 9.00260197203842689217E1,
                                2.23200534594684319226E3, 7.00332514112805075473E3,
                                5.55923013010394962768E4};
    static const double U[5] = {3.35617141647503099647E1, 5.21357949780152679795E2,
                                4.59432382970980127987E3, 2.26290000613890934246E4,
                                4.92673942608635921086E4};
    double z = x * x;
    double t = T[0];
    for (int i = 1; i < 5; ++i) t = t * z + T[i];
    double u = z + U[0];
    for (int i = 1; i < 5; ++i) u = u * z + U[i];
    return x * t / u;
}

double orc_erf(double x) { return orc_erf_ref(x); }

static inline void g_dgdr(int kernel, double rb, double *g, double *dgdr)
{
    if (kernel == 1) {
        double r2 = rb * rb;
        *g = rb * rb * rb * (r2 + 2.5) / pow(r2 + 1.0, 2.5);
        *dgdr = 7.5 * r2 / pow(r2 + 1.0, 3.5);
    } else {
        const double c = 0.7978845608028654;          /* sqrt(2/pi) */
        double ev = exp(-0.5 * rb * rb);
        *g = orc_erf_ref(rb / 1.4142135623730951) - c * rb * ev;
        *dgdr = c * rb * rb * ev;
    }
}

/* u(x) = -(1/4pi) sum_j g(r/sigma_j)/r^3 (x - x_j) x alpha_j   (kernel.py:103-114).
 * The self pair (r < 1e-12) is excluded as in warp_vpm.py:37-38; for the
 * velocity this is identical to kernel.py's clamp because dx = 0 there. */
int orc_particle_vel(const double *tgt, int64_t T, const double *pos, const double *alpha,
                     const double *sigma, int64_t N, int kernel, double *out, int nthreads)
{
    return orc_particle_vel_rule(tgt, T, pos, alpha, sigma, N, kernel, 0, out, nthreads);
}

/* rule 0: pairs closer than 1e-12 are skipped (warp_vpm.py:37-38);
 * rule 1: r = max(r, 1e-12) (kernel.py:105-106), the reference CPU path. */
int orc_particle_vel_rule(const double *tgt, int64_t T, const double *pos, const double *alpha,
                          const double *sigma, int64_t N, int kernel, int rule, double *out,
                          int nthreads)
{
    const double C = -1.0 / (4.0 * M_PI);
    memset(out, 0, sizeof(double) * 3 * (size_t)T);
    if (N == 0) return 0;
#pragma omp parallel for schedule(static) num_threads(orc_threads(nthreads))
    for (int64_t i = 0; i < T; ++i) {
        double ux = 0.0, uy = 0.0, uz = 0.0;
        const double *x = tgt + 3 * i;
        for (int64_t j = 0; j < N; ++j) {
            double dx = x[0] - pos[3 * j], dy = x[1] - pos[3 * j + 1], dz = x[2] - pos[3 * j + 2];
            double r2 = dx * dx + dy * dy + dz * dz;
            double r = sqrt(r2);
            if (r < 1.0e-12) {
                if (rule == 0) continue;
                r = 1.0e-12;
            }
            double rb = r / sigma[j];
            double g, dg;
            g_dgdr(kernel, rb, &g, &dg);
            const double *a = alpha + 3 * j;
            double cx = dy * a[2] - dz * a[1];
            double cy = dz * a[0] - dx * a[2];
            double cz = dx * a[1] - dy * a[0];
            double coeff = C * g / (r * r * r);
            ux += coeff * cx;
            uy += coeff * cy;
            uz += coeff * cz;
        }
        out[3 * i] = ux;
        out[3 * i + 1] = uy;
        out[3 * i + 2] = uz;
    }
    return 0;
}

/* J[t][i][j] = dU_i/dx_j, kernel.py:156-188, self pair skipped (warp_vpm.py:103). */
int orc_particle_jac(const double *tgt, int64_t T, const double *pos, const double *alpha,
                     const double *sigma, int64_t N, int kernel, double *J, int nthreads)
{
    return orc_particle_jac_rule(tgt, T, pos, alpha, sigma, N, kernel, 0, J, nthreads);
}

/* rule 0: pairs closer than 1e-12 skipped (warp_vpm.py:103); rule 1:
 * r = sqrt(max(r^2, 1e-24)) (kernel.py:156), which keeps the self pair and
 * exactly coincident pairs: their dx = 0 terms vanish, the aux2 terms
 * (aux2 = -g(1e-12/sigma) / (4 pi 1e-36), ~1e8 |alpha|) do not.  Those
 * clamped pairs are summed apart and added last, so coincident particles of
 * opposite alpha cancel exactly; the reference sums them among the others,
 * where they leave round-off of their own size (~1e-8 |alpha|) that depends
 * on NumPy's summation order. */
int orc_particle_jac_rule(const double *tgt, int64_t T, const double *pos, const double *alpha,
                          const double *sigma, int64_t N, int kernel, int rule, double *J,
                          int nthreads)
{
    const double C = -1.0 / (4.0 * M_PI);
    memset(J, 0, sizeof(double) * 9 * (size_t)T);
    if (N == 0) return 0;
#pragma omp parallel for schedule(static) num_threads(orc_threads(nthreads))
    for (int64_t t = 0; t < T; ++t) {
        double main_acc[9] = {0}, clamped[9] = {0};
        const double *x = tgt + 3 * t;
        for (int64_t j = 0; j < N; ++j) {
            double d[3] = {x[0] - pos[3 * j], x[1] - pos[3 * j + 1], x[2] - pos[3 * j + 2]};
            double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
            double r;
            double *acc = main_acc;
            if (rule == 0) {
                r = sqrt(r2);
                if (r < 1.0e-12) continue;
            } else {
                r = sqrt(r2 > 1.0e-24 ? r2 : 1.0e-24);
                if (r2 < 1.0e-24) acc = clamped;
            }
            double s = sigma[j];
            double rb = r / s;
            double g, dg;
            g_dgdr(kernel, rb, &g, &dg);
            const double *a = alpha + 3 * j;
            double r3inv = 1.0 / (r * r * r);
            double k[3] = {C * (d[1] * a[2] - d[2] * a[1]) * r3inv,
                           C * (d[2] * a[0] - d[0] * a[2]) * r3inv,
                           C * (d[0] * a[1] - d[1] * a[0]) * r3inv};
            double aux = dg / (s * r) - 3.0 * g / (r * r);
            double aux2 = C * g * r3inv;
            acc[0] += aux * k[0] * d[0];
            acc[1] += aux * k[0] * d[1] + aux2 * a[2];
            acc[2] += aux * k[0] * d[2] - aux2 * a[1];
            acc[3] += aux * k[1] * d[0] - aux2 * a[2];
            acc[4] += aux * k[1] * d[1];
            acc[5] += aux * k[1] * d[2] + aux2 * a[0];
            acc[6] += aux * k[2] * d[0] + aux2 * a[1];
            acc[7] += aux * k[2] * d[1] - aux2 * a[0];
            acc[8] += aux * k[2] * d[2];
        }
        for (int c = 0; c < 9; ++c) J[9 * t + c] = main_acc[c] + clamped[c];
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* VortexParticleField.advect_rk3, particles.py:85-199.                     */
/* Constant freestream U_inf (HybridSolver passes a broadcast vector,       */
/* experiment_hybrid_panel_particle.py:140-143); no bound-velocity callback. */
/* ------------------------------------------------------------------------ */
int orc_vpm_advect_rk3(int64_t n, double *pos, double *gamma, double *sigma, double *age,
                       double dt, const double *Uinf, int stretch, int free_wake, int kernel,
                       double nu, int nthreads)
{
    return orc_vpm_advect_rk3_rule(n, pos, gamma, sigma, age, dt, Uinf, stretch, free_wake,
                                   kernel, nu, 0, nthreads);
}

static int advect_impl(int64_t n, double *pos, double *gamma, double *sigma, double *age,
                       double dt, const double *Uinf, orc_stage_velocity_fn ext, void *user,
                       int stretch, int free_wake, int kernel, double nu, int rule,
                       int nthreads);

int orc_vpm_advect_rk3_rule(int64_t n, double *pos, double *gamma, double *sigma, double *age,
                            double dt, const double *Uinf, int stretch, int free_wake,
                            int kernel, double nu, int rule, int nthreads)
{
    return advect_impl(n, pos, gamma, sigma, age, dt, Uinf, NULL, NULL, stretch, free_wake,
                       kernel, nu, rule, nthreads);
}

/* advect_rk3 with U_inf_func and bound_velocity_func evaluated at every stage's
 * positions (particles.py:126-133): ext fills U_bound and U_inf (n, 3). */
int orc_vpm_advect_rk3_ext(int64_t n, double *pos, double *gamma, double *sigma, double *age,
                           double dt, orc_stage_velocity_fn ext, void *user, int stretch,
                           int free_wake, int kernel, double nu, int rule, int nthreads)
{
    if (!ext) return -1;
    return advect_impl(n, pos, gamma, sigma, age, dt, NULL, ext, user, stretch, free_wake,
                       kernel, nu, rule, nthreads);
}

static int advect_impl(int64_t n, double *pos, double *gamma, double *sigma, double *age,
                       double dt, const double *Uinf, orc_stage_velocity_fn ext, void *user,
                       int stretch, int free_wake, int kernel, double nu, int rule,
                       int nthreads)
{
    if (n == 0) return 0;
    static const double ra[3] = {0.0, -5.0 / 9.0, -153.0 / 128.0};   /* :114 */
    static const double rb[3] = {1.0 / 3.0, 15.0 / 16.0, 8.0 / 15.0};
    const double max_gamma_sq = 1e4;                                 /* :117 */
    size_t n3 = 3 * (size_t)n;
    double *rkX = calloc(n3, sizeof(double));
    double *rkG = calloc(n3, sizeof(double));
    double *rks = calloc((size_t)n, sizeof(double));
    double *U = calloc(n3, sizeof(double));
    double *J = stretch ? calloc(9 * (size_t)n, sizeof(double)) : NULL;
    double *Ub = ext ? calloc(n3, sizeof(double)) : NULL;
    double *Ui = ext ? calloc(n3, sizeof(double)) : NULL;
    if (!rkX || !rkG || !rks || !U || (stretch && !J) || (ext && (!Ub || !Ui))) {
        free(rkX); free(rkG); free(rks); free(U); free(J); free(Ub); free(Ui);
        return -1;
    }
    for (int st = 0; st < 3; ++st) {
        double a = ra[st], b = rb[st];
        if (free_wake)
            orc_particle_vel_rule(pos, n, pos, gamma, sigma, n, kernel, rule, U, nthreads);
        else
            memset(U, 0, n3 * sizeof(double));
        if (ext && ext(user, n, pos, Ub, Ui) != 0) {                      /* :126-131 */
            free(rkX); free(rkG); free(rks); free(U); free(J); free(Ub); free(Ui);
            return -2;
        }
        for (int64_t i = 0; i < n; ++i) {
            /* U_total = U_part + U_bound + U_inf, :133 */
            double u[3];
            for (int c = 0; c < 3; ++c)
                u[c] = ext ? (U[3 * i + c] + Ub[3 * i + c]) + Ui[3 * i + c]
                           : U[3 * i + c] + 0.0 + Uinf[c];
            double sp = sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);   /* :136-140 */
            if (sp > 50.0) {
                double f = 50.0 / sp;
                for (int c = 0; c < 3; ++c) u[c] *= f;
            }
            for (int c = 0; c < 3; ++c) {                               /* :143-144 */
                rkX[3 * i + c] = a * rkX[3 * i + c] + dt * u[c];
                pos[3 * i + c] += b * rkX[3 * i + c];
            }
        }
        if (stretch) {
            orc_particle_jac_rule(pos, n, pos, gamma, sigma, n, kernel, rule, J,
                                  nthreads);                                    /* :148 */
            double zmax = 10.0 / (dt > 1e-10 ? dt : 1e-10);                     /* :161 */
            for (int64_t i = 0; i < n; ++i) {
                const double *Ji = J + 9 * i;
                double *g = gamma + 3 * i;
                double S[3];
                for (int r = 0; r < 3; ++r)                                     /* :151 */
                    S[r] = Ji[3 * r] * g[0] + Ji[3 * r + 1] * g[1] + Ji[3 * r + 2] * g[2];
                double gsq = g[0] * g[0] + g[1] * g[1] + g[2] * g[2];
                double sdg = S[0] * g[0] + S[1] * g[1] + S[2] * g[2];
                double Z = 0.0;
                if (gsq > 1e-20) Z = 0.2 * sdg / gsq;                            /* :157-159 */
                if (Z < -zmax) Z = -zmax;
                if (Z > zmax) Z = zmax;
                double ds = -sigma[i] * Z;                                       /* :165 */
                if (nu > 0) ds += nu / (sigma[i] > 1e-10 ? sigma[i] : 1e-10);
                for (int c = 0; c < 3; ++c) {
                    double dG = S[c] - 3.0 * Z * g[c];
                    rkG[3 * i + c] = a * rkG[3 * i + c] + dt * dG;
                }
                rks[i] = a * rks[i] + dt * ds;
                for (int c = 0; c < 3; ++c) g[c] += b * rkG[3 * i + c];
                sigma[i] += b * rks[i];
                double g2 = g[0] * g[0] + g[1] * g[1] + g[2] * g[2];            /* :177-181 */
                if (g2 > max_gamma_sq) {
                    double sc = sqrt(max_gamma_sq / g2);
                    for (int c = 0; c < 3; ++c) g[c] *= sc;
                }
                if (sigma[i] < 1e-6) sigma[i] = 1e-6;                            /* :183 */
                if (sigma[i] > 5.0) sigma[i] = 5.0;
            }
        }
    }
    for (int64_t i = 0; i < n; ++i) {                                            /* :186-196 */
        int bad = 0;
        for (int c = 0; c < 3; ++c)
            if (!isfinite(pos[3 * i + c]) || !isfinite(gamma[3 * i + c])) bad = 1;
        if (bad) gamma[3 * i] = gamma[3 * i + 1] = gamma[3 * i + 2] = 0.0;
        age[i] += dt;
    }
    free(rkX); free(rkG); free(rks); free(U); free(J); free(Ub); free(Ui);
    return 0;
}

/* Pedrizzetti relaxation, particles.py:202-246.  Always uses the Gaussian-erf
 * Jacobian (the reference calls jacobian_from_particles without kernel=, :219). */
int orc_vpm_relax(int64_t n, const double *pos, double *gamma, const double *sigma,
                  double rlxf, int nthreads)
{
    return orc_vpm_relax_rule(n, pos, gamma, sigma, rlxf, 0, nthreads);
}

int orc_vpm_relax_rule(int64_t n, const double *pos, double *gamma, const double *sigma,
                       double rlxf, int rule, int nthreads)
{
    if (rlxf <= 0 || n == 0) return 0;
    int any = 0;
    for (int64_t i = 0; i < n && !any; ++i) {
        const double *g = gamma + 3 * i;
        if (sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]) > 1e-15) any = 1;
    }
    if (!any) return 0;                                   /* :214-216 */
    double *J = calloc(9 * (size_t)n, sizeof(double));
    if (!J) return -1;
    orc_particle_jac_rule(pos, n, pos, gamma, sigma, n, 0, rule, J, nthreads);
    for (int64_t i = 0; i < n; ++i) {
        double *g = gamma + 3 * i;
        const double *Ji = J + 9 * i;
        double gn = sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]);
        double w[3] = {Ji[7] - Ji[5], Ji[2] - Ji[6], Ji[3] - Ji[1]};   /* :223-225 */
        double wn = sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
        if (!(gn > 1e-15 && wn > 1e-12)) continue;                     /* :228 */
        double wh[3] = {w[0] / wn, w[1] / wn, w[2] / wn};
        for (int c = 0; c < 3; ++c) g[c] = (1.0 - rlxf) * g[c] + rlxf * gn * wh[c];
        double dot = g[0] * wh[0] + g[1] * wh[1] + g[2] * wh[2];
        double b2 = 1.0 - 2.0 * (1.0 - rlxf) * rlxf * (1.0 - dot / (gn > 1e-12 ? gn : 1e-12));
        if (b2 < 0.01) b2 = 0.01;                                       /* :243 */
        double sb = sqrt(b2);
        for (int c = 0; c < 3; ++c) g[c] /= sb;
    }
    free(J);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Panel -> particle conversion, experiment_hybrid_panel_particle.py:173-200. */
/* Rings are visited in the given order; each ring with |Gamma| >= 1e-15     */
/* yields up to 4 particles (front, left, back, right legs) whose            */
/* |alpha|^2 > 1e-30, and its strength is set to 0.  Returns the count.      */
/* ------------------------------------------------------------------------ */
int64_t orc_convert_rings(int64_t nr, const double *Br, const double *Fr, const double *Fl,
                          const double *Bl, double *gamma, double sigma, double *pos_out,
                          double *alpha_out, double *sig_out)
{
    int64_t k = 0;
    for (int64_t r = 0; r < nr; ++r) {
        double G = gamma[r];
        if (fabs(G) < 1e-15) continue;                    /* :179 */
        const double *fr = Fr + 3 * r, *fl = Fl + 3 * r, *bl = Bl + 3 * r, *br = Br + 3 * r;
        const double *ends[4][2] = {{fr, fl}, {fl, bl}, {bl, br}, {br, fr}};   /* :187-192 */
        for (int l = 0; l < 4; ++l) {
            const double *a = ends[l][0], *b = ends[l][1];
            double al[3], ps[3];
            for (int c = 0; c < 3; ++c) {
                ps[c] = 0.5 * (a[c] + b[c]);
                al[c] = G * (b[c] - a[c]);
            }
            if (al[0] * al[0] + al[1] * al[1] + al[2] * al[2] > 1e-30) {   /* :195 */
                memcpy(pos_out + 3 * k, ps, sizeof ps);
                memcpy(alpha_out + 3 * k, al, sizeof al);
                sig_out[k] = sigma;
                ++k;
            }
        }
        gamma[r] = 0.0;                                    /* :200 */
    }
    return k;
}
