/* fv_oracle.h -- declarations of the CPU FP64 oracle (test infrastructure
 * only; see the header comment of fv_oracle.c for who may call it). */
#ifndef FV_ORACLE_H
#define FV_ORACLE_H
#include <stdint.h>

int orc_collapsed(const double *P, int64_t N, const double *Br, const double *Fr,
                  const double *Fl, const double *Bl, int64_t M, const double *gamma,
                  const double *rc0, const double *age, double nu, int leg_mask, double *out,
                  int nthreads);
int orc_expanded(const double *P, int64_t N, const double *Br, const double *Fr,
                 const double *Fl, const double *Bl, int64_t M, const double *gamma,
                 const double *rc0, const double *age, double nu, int leg_mask, double *out,
                 int nthreads);
int orc_aic(const double *colloc, const double *normals, int64_t P, const double *Br,
            const double *Fr, const double *Fl, const double *Bl, const double *rc0, double nu,
            double *aic, int nthreads);
int orc_lu_solve(int64_t n, double *A, double *b);
int orc_particle_vel(const double *tgt, int64_t T, const double *pos, const double *alpha,
                     const double *sigma, int64_t N, int kernel, double *out, int nthreads);
int orc_particle_jac(const double *tgt, int64_t T, const double *pos, const double *alpha,
                     const double *sigma, int64_t N, int kernel, double *J, int nthreads);
int orc_vpm_advect_rk3(int64_t n, double *pos, double *gamma, double *sigma, double *age,
                       double dt, const double *Uinf, int stretch, int free_wake, int kernel,
                       double nu, int nthreads);
int orc_vpm_relax(int64_t n, const double *pos, double *gamma, const double *sigma,
                  double rlxf, int nthreads);
/* Pair rule for particles closer than 1e-12: 0 skips them (warp_vpm.py:37,103),
 * 1 clamps r to 1e-12 as the reference's CPU kernels do (kernel.py:105,156). */
int orc_particle_vel_rule(const double *tgt, int64_t T, const double *pos, const double *alpha,
                          const double *sigma, int64_t N, int kernel, int rule, double *out,
                          int nthreads);
int orc_particle_jac_rule(const double *tgt, int64_t T, const double *pos, const double *alpha,
                          const double *sigma, int64_t N, int kernel, int rule, double *J,
                          int nthreads);
int orc_vpm_advect_rk3_rule(int64_t n, double *pos, double *gamma, double *sigma, double *age,
                            double dt, const double *Uinf, int stretch, int free_wake,
                            int kernel, double nu, int rule, int nthreads);
int orc_vpm_relax_rule(int64_t n, const double *pos, double *gamma, const double *sigma,
                       double rlxf, int rule, int nthreads);
/* Stage velocity of advect_rk3's callables (particles.py:126-131): the stage
 * positions (n, 3) in, U_bound and U_inf (n, 3) out; nonzero aborts. */
typedef int (*orc_stage_velocity_fn)(void *user, int64_t n, const double *pos, double *u_bound,
                                     double *u_inf);
int orc_vpm_advect_rk3_ext(int64_t n, double *pos, double *gamma, double *sigma, double *age,
                           double dt, orc_stage_velocity_fn ext, void *user, int stretch,
                           int free_wake, int kernel, double nu, int rule, int nthreads);
double orc_erf(double x);
int64_t orc_convert_rings(int64_t nr, const double *Br, const double *Fr, const double *Fl,
                          const double *Bl, double *gamma, double sigma, double *pos_out,
                          double *alpha_out, double *sig_out);
int64_t orc_pop_drop(int64_t np, int64_t max_particles, double *pos, double *alpha,
                     double *sigma, double *age);
int64_t orc_pop_merge(int64_t np, double *pos, double *alpha, double *sigma, double *age,
                      const double *ref, int64_t nref, double eps, double vinf_norm,
                      int64_t protect, double protect_dist, int64_t *n_merged);
void orc_ledger(int64_t np, const double *alpha, int64_t nr, const double *Br, const double *Fr,
                const double *Fl, const double *Bl, const double *gamma, double *out);

#endif
