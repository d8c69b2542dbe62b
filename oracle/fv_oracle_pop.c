/* fv_oracle_pop.c -- CPU FP64 restatement of the particle population control
 * and the vector-circulation ledger (test infrastructure only: the tests,
 * the smoke check and bench.py's CPU baseline may call it; the product never
 * does).
 *
 * Restated from newton_pc/adapters/flap.py:
 *   drop    :469-473  after convection, beyond max_particles the oldest
 *                     particles (the front of the arrays) are removed;
 *   merge   :478-563  pairwise moment-conserving merging with an at-wing
 *                     velocity error criterion;
 *   ledger  :565-578  sum of particle alpha + sum over ring legs of G (b - a).
 *
 * Merge, step by step (the device follows the same statement):
 *   1. P < 2 protect: nothing.  Wing centre c = (sum of the reference points
 *      in order) / n_ref.
 *   2. Candidates: with protect_dist > 0 the particles with |x - c| >
 *      protect_dist (fewer than 4: nothing); else the oldest P - protect.
 *   3. Per candidate: d = |x - c|, cell = 0.1 + 0.15 d, key_k = floor(x_k /
 *      cell) (k = 0, 1, 2).
 *   4. Stable sort of the candidates by (key0, key1, key2).
 *   5. Greedy scan of the sorted order: consecutive candidates i, j with equal
 *      keys, both unused and alpha_i . alpha_j > 0, form a pair.
 *   6. Per pair: am = a1 + a2, xm = (|a1| x1 + |a2| x2) / (|a1| + |a2| + 1e-30),
 *      sm = max(s1, s2); dv = v(x1, a1, s1) + v(x2, a2, s2) - v(xm, am, sm) at
 *      every reference point with v = g(rho) / (4 pi d^3) a x r, r = R - X,
 *      d = |r| + 1e-30, rho = d / s, g = rho^3 / (rho^2 + 1)^1.5; accepted
 *      when max |dv| < eps |V_inf|.
 *   7. Accepted pairs: the first particle takes (xm, am, sm) (its age is
 *      kept; the reference's adapter carries no age), the second is removed;
 *      the survivors keep their order.
 * Arithmetic: plain double operations in the order written (-ffp-contract=off),
 * d^3 as d d d, rho^3 as rho rho rho and (rho^2 + 1)^1.5 as t sqrt(t). */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "fv_oracle.h"

/* flap.py:469-473 */
int64_t orc_pop_drop(int64_t np, int64_t max_particles, double *pos, double *alpha,
                     double *sigma, double *age)
{
    if (np <= max_particles) return np;
    const int64_t drop = np - max_particles;
    memmove(pos, pos + 3 * drop, sizeof(double) * 3 * (size_t)max_particles);
    memmove(alpha, alpha + 3 * drop, sizeof(double) * 3 * (size_t)max_particles);
    memmove(sigma, sigma + drop, sizeof(double) * (size_t)max_particles);
    if (age) memmove(age, age + drop, sizeof(double) * (size_t)max_particles);
    return max_particles;
}

typedef struct {
    int64_t k0, k1, k2;
    int64_t idx; /* candidate ordinal: ties keep the original order */
} key_t_;

static int key_cmp(const void *a, const void *b)
{
    const key_t_ *x = a, *y = b;
    if (x->k0 != y->k0) return x->k0 < y->k0 ? -1 : 1;
    if (x->k1 != y->k1) return x->k1 < y->k1 ? -1 : 1;
    if (x->k2 != y->k2) return x->k2 < y->k2 ? -1 : 1;
    return x->idx < y->idx ? -1 : (x->idx > y->idx);
}

/* v(R; X, A, S) accumulated into v with sign sg (flap.py:536-543) */
static void merge_vel(const double *R, const double *X, const double *A, double S, double *v)
{
    const double r[3] = {R[0] - X[0], R[1] - X[1], R[2] - X[2]};
    const double d = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]) + 1e-30;
    const double rho = d / S;
    const double t = rho * rho + 1.0;
    const double g = rho * rho * rho / (t * sqrt(t));
    const double c = g / (4.0 * M_PI * (d * d * d));
    v[0] = c * (A[1] * r[2] - A[2] * r[1]);
    v[1] = c * (A[2] * r[0] - A[0] * r[2]);
    v[2] = c * (A[0] * r[1] - A[1] * r[0]);
}

/* flap.py:478-563.  Returns the new count (n_merged set), -1 on allocation
 * failure. */
int64_t orc_pop_merge(int64_t np, double *pos, double *alpha, double *sigma, double *age,
                      const double *ref, int64_t nref, double eps, double vinf_norm,
                      int64_t protect, double protect_dist, int64_t *n_merged)
{
    *n_merged = 0;
    if (np < 2 * protect || nref < 1) return np;
    double c[3] = {0.0, 0.0, 0.0};
    for (int64_t k = 0; k < nref; ++k)
        for (int q = 0; q < 3; ++q) c[q] += ref[3 * k + q];
    for (int q = 0; q < 3; ++q) c[q] /= (double)nref;
    int64_t *cand = malloc(sizeof(int64_t) * (size_t)(np > 0 ? np : 1));
    key_t_ *keys = malloc(sizeof(key_t_) * (size_t)(np > 0 ? np : 1));
    unsigned char *used = calloc((size_t)(np > 0 ? np : 1), 1);
    unsigned char *keep = malloc((size_t)(np > 0 ? np : 1));
    int64_t *pi = malloc(sizeof(int64_t) * (size_t)(np / 2 + 1));
    int64_t *pj = malloc(sizeof(int64_t) * (size_t)(np / 2 + 1));
    if (!cand || !keys || !used || !keep || !pi || !pj) {
        free(cand); free(keys); free(used); free(keep); free(pi); free(pj);
        return -1;
    }
    int64_t nf = 0;
    for (int64_t k = 0; k < np; ++k) {
        const double *x = pos + 3 * k;
        const double dx = x[0] - c[0], dy = x[1] - c[1], dz = x[2] - c[2];
        const double d = sqrt(dx * dx + dy * dy + dz * dz);
        if (protect_dist > 0.0 ? d > protect_dist : k < np - protect) cand[nf++] = k;
    }
    int64_t result = np;
    if (protect_dist > 0.0 && nf < 4) goto out;
    for (int64_t a = 0; a < nf; ++a) {
        const double *x = pos + 3 * cand[a];
        const double dx = x[0] - c[0], dy = x[1] - c[1], dz = x[2] - c[2];
        const double d = sqrt(dx * dx + dy * dy + dz * dz);
        const double cell = 0.1 + 0.15 * d;
        keys[a].k0 = (int64_t)floor(x[0] / cell);
        keys[a].k1 = (int64_t)floor(x[1] / cell);
        keys[a].k2 = (int64_t)floor(x[2] / cell);
        keys[a].idx = a;
    }
    qsort(keys, (size_t)nf, sizeof(key_t_), key_cmp);
    int64_t npair = 0;
    for (int64_t a = 0; a + 1 < nf; ++a) {
        if (keys[a].k0 != keys[a + 1].k0 || keys[a].k1 != keys[a + 1].k1 ||
            keys[a].k2 != keys[a + 1].k2)
            continue;
        const int64_t i = keys[a].idx, j = keys[a + 1].idx;
        if (used[i] || used[j]) continue;
        const double *ai = alpha + 3 * cand[i], *aj = alpha + 3 * cand[j];
        if (ai[0] * aj[0] + ai[1] * aj[1] + ai[2] * aj[2] <= 0.0) continue;
        used[i] = used[j] = 1;
        pi[npair] = cand[i];
        pj[npair] = cand[j];
        ++npair;
    }
    memset(keep, 1, (size_t)np);
    const double thr = eps * vinf_norm;
    int64_t nm = 0;
    for (int64_t q = 0; q < npair; ++q) {
        const int64_t i = pi[q], j = pj[q];
        const double *a1 = alpha + 3 * i, *a2 = alpha + 3 * j;
        const double *x1 = pos + 3 * i, *x2 = pos + 3 * j;
        const double w1 = sqrt(a1[0] * a1[0] + a1[1] * a1[1] + a1[2] * a1[2]);
        const double w2 = sqrt(a2[0] * a2[0] + a2[1] * a2[1] + a2[2] * a2[2]);
        double am[3], xm[3];
        for (int k = 0; k < 3; ++k) {
            am[k] = a1[k] + a2[k];
            xm[k] = (w1 * x1[k] + w2 * x2[k]) / (w1 + w2 + 1e-30);
        }
        const double sm = sigma[i] > sigma[j] ? sigma[i] : sigma[j];
        double err = 0.0;
        for (int64_t r = 0; r < nref; ++r) {
            double v1[3], v2[3], v3[3];
            merge_vel(ref + 3 * r, x1, a1, sigma[i], v1);
            merge_vel(ref + 3 * r, x2, a2, sigma[j], v2);
            merge_vel(ref + 3 * r, xm, am, sm, v3);
            for (int k = 0; k < 3; ++k) {
                const double dv = fabs((v1[k] + v2[k]) - v3[k]);
                if (dv > err) err = dv;
            }
        }
        if (!(err < thr)) continue;
        memcpy(pos + 3 * i, xm, sizeof xm);
        memcpy(alpha + 3 * i, am, sizeof am);
        sigma[i] = sm;
        keep[j] = 0;
        ++nm;
    }
    int64_t w = 0;
    for (int64_t k = 0; k < np; ++k) {
        if (!keep[k]) continue;
        if (w != k) {
            memcpy(pos + 3 * w, pos + 3 * k, 3 * sizeof(double));
            memcpy(alpha + 3 * w, alpha + 3 * k, 3 * sizeof(double));
            sigma[w] = sigma[k];
            if (age) age[w] = age[k];
        }
        ++w;
    }
    *n_merged = nm;
    result = w;
out:
    free(cand); free(keys); free(used); free(keep); free(pi); free(pj);
    return result;
}

/* flap.py:565-578: sum of particle alpha (in order) + sum over rings of
 * G (b - a) over the legs Fl->Fr, Fr->Br, Br->Bl, Bl->Fl of each ring. */
void orc_ledger(int64_t np, const double *alpha, int64_t nr, const double *Br, const double *Fr,
                const double *Fl, const double *Bl, const double *gamma, double *out)
{
    double t[3] = {0.0, 0.0, 0.0};
    for (int64_t k = 0; k < np; ++k)
        for (int q = 0; q < 3; ++q) t[q] += alpha[3 * k + q];
    for (int64_t k = 0; k < nr; ++k) {
        const double *cc[4] = {Fl + 3 * k, Fr + 3 * k, Br + 3 * k, Bl + 3 * k};
        for (int l = 0; l < 4; ++l) {
            const double *a = cc[l], *b = cc[(l + 1) & 3];
            for (int q = 0; q < 3; ++q) t[q] = t[q] + gamma[k] * (b[q] - a[q]);
        }
    }
    memcpy(out, t, sizeof t);
}
