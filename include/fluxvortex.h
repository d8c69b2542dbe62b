/*
 * fluxvortex.h -- C-ABI of libfluxvortex_hip.so, the MI355X (gfx950) hot path
 * of FLUXVortex's hybrid panel-particle solver.
 *
 * Conventions
 *   - Plain pointers and sizes only.  "_host" entry points take host arrays,
 *     copy them to device buffers owned by the context, run the kernels on the
 *     context's stream and copy the result back before returning (blocking,
 *     like the reference's D2H copy at warp_kernels.py:214).  "_dev" entry
 *     points take device pointers, enqueue on the context stream and return
 *     without synchronising.
 *   - Point/vertex arrays are C-contiguous AoS (n, 3) FP64, as the reference's
 *     NumPy arrays (warp_kernels.py:251-262).
 *   - Every function returns 0 on success or a negative FV_E* code; the
 *     message is available from fv_last_error(ctx).  No C++ exception crosses
 *     this boundary.
 *   - A context owns one HIP stream and grow-only device workspaces; use one
 *     context per host thread.
 *
 * Reference paths below are relative to the FLUXVortex repository.
 */
#ifndef FLUXVORTEX_H
#define FLUXVORTEX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FV_OK 0
#define FV_E_ARG (-1)     /* invalid argument (null pointer, negative size, bad enum) */
#define FV_E_HIP (-2)     /* HIP runtime error (allocation, copy, launch) */
#define FV_E_CAP (-3)     /* capacity exceeded (particle field full) */
#define FV_E_SING (-4)    /* singular matrix in the dense solve */

/* Particle smoothing kernels, kernel.py:59-62. */
#define FV_KERNEL_GAUSSIAN_ERF 0
#define FV_KERNEL_WINCKELMANS 1

/* Arithmetic of the pair sums. */
#define FV_PREC_FP64 0    /* plain FP64 */
#define FV_PREC_MIXED 1   /* FP32 pair terms (segments) / kernel factors (particles),
                            * FP64 sums; FP32 LU + FP64-residual refinement */

/* Particles closer than 1e-12 (the self pair, exactly coincident particles). */
#define FV_PAIRS_SKIP_CLOSE 0   /* skipped, as the reference's GPU twin warp_vpm.py:37,103 */
#define FV_PAIRS_CLAMP 1        /* r clamped to 1e-12, as the reference's CPU field
                                 * (kernel.py:105,156) that HybridSolver runs */

/* Leg masks over ring corners (Br, Fr, Fl, Bl); leg k runs corner k -> k+1.
 * warp_kernels.py:268-273 (ring), :296-299 (chordwise), :350-354 (horseshoe). */
#define FV_LEGS_RING 0xF
#define FV_LEGS_CHORDWISE 0x5
#define FV_LEGS_HORSESHOE 0x7

typedef struct fv_ctx fv_ctx;
typedef struct fv_vpm fv_vpm;
typedef struct fv_hybrid fv_hybrid;
typedef struct fv_hybrid_batch fv_hybrid_batch;

/* ---------------------------------------------------------------- context */
/* No reference counterpart (Warp keeps an implicit device context,
 * warp_fsi/config.py:60-63); owns the stream and workspaces. */
int fv_ctx_create(int device, fv_ctx **out);
void fv_ctx_destroy(fv_ctx *ctx);
const char *fv_last_error(const fv_ctx *ctx);
int fv_ctx_synchronize(fv_ctx *ctx);
void *fv_ctx_stream(fv_ctx *ctx);
int fv_device_alloc(fv_ctx *ctx, size_t bytes, void **out);
int fv_device_free(fv_ctx *ctx, void *ptr);
int fv_memcpy_h2d(fv_ctx *ctx, void *dst, const void *src, size_t bytes);
int fv_memcpy_d2h(fv_ctx *ctx, void *dst, const void *src, size_t bytes);
/* Per-kernel-family timing with HIP events on the context stream.  While
 * enabled, every launch of a timed family ("seg_pairs", "particle_pairs",
 * "aic", "dense_solve", ...) is bracketed by events; fv_ctx_kernel_stats
 * synchronises and returns the launch count and summed duration.  Enabling
 * clears earlier records. */
int fv_ctx_set_timing(fv_ctx *ctx, int enabled);
/* Project-written autotuner of the pair-sum tile shapes (64 / 256 lanes,
 * 1 / 2 / 4 targets per lane, 4 / 8 / 16 workgroups per CU): when enabled,
 * the first pair sum of each kind (operator, device, power-of-two size class
 * of targets and sources) times every shape with HIP events (synchronising
 * once) and every later call of that kind uses the fastest.  Shapes only
 * change how sources are split, so results differ at round-off level from
 * the default shape.  Sums beyond 2^36 pairs are not tuned (36 trial sums
 * would take seconds); they keep the default shape.  Off by default (no
 * counterpart in the reference). */
int fv_ctx_set_autotune(fv_ctx *ctx, int32_t enabled);
/* The tuned kinds, one text line each: family log2_targets log2_sources
 * lanes targets_per_lane workgroups_per_cu best_ms. */
int fv_autotune_report(fv_ctx *ctx, char *buf, int64_t cap);
int fv_ctx_kernel_stats(fv_ctx *ctx, const char *family, int64_t *launches, double *total_ms);
/* Eight user events on the context stream (which = 0..7). */
int fv_ctx_event_record(fv_ctx *ctx, int32_t which);
int fv_ctx_event_elapsed(fv_ctx *ctx, int32_t a, int32_t b, double *ms);

/* ------------------------------------------- line-vortex Biot-Savart (a1-a5) */
/* Collapsed sum over vortices and the legs in leg_mask:
 *   out[i] = sum_m sum_legs v(points[i]; leg of vortex m)            (n, 3)
 * Replaces warp_kernels.py:251-276 (ring), :279-302 (chordwise), :335-357
 * (horseshoe; pass ages = NULL), i.e. PteraSoftware's
 * _aerodynamics_functions.collapsed_velocities_from_* .  ages == NULL means
 * zeros (:265-266).  n == 0 or m == 0 gives zeros (:193-194). */
int fv_vortex_collapsed_host(fv_ctx *ctx, const double *points, int64_t n, const double *br,
                             const double *fr, const double *fl, const double *bl, int64_t m,
                             const double *strengths, const double *r_c0s, const double *ages,
                             double nu, int32_t leg_mask, int32_t precision, double *out);
int fv_vortex_collapsed_dev(fv_ctx *ctx, const double *points, int64_t n, const double *br,
                            const double *fr, const double *fl, const double *bl, int64_t m,
                            const double *strengths, const double *r_c0s, const double *ages,
                            double nu, int32_t leg_mask, int32_t precision, double *out);
/* fv_vortex_collapsed_dev without any allocation.  No counterpart in the
 * reference: this entry point is for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  Same arguments and results as fv_vortex_collapsed_dev
 * with the autotuner off (the source split may be capped, which moves the
 * result by round-off only).  Its segment table and partial sums live in a
 * 32 MiB file-scope device array that every GPU the library runs on holds
 * until the process ends: at most 419 430 segments (m x legs); beyond that
 * FV_E_ARG with a message.  Calls that overlap on one GPU (several contexts
 * or streams at once) race on that array; calls on one stream are ordered. */
int fv_vortex_collapsed_dev_noalloc(fv_ctx *ctx, const double *points, int64_t n,
                                    const double *br, const double *fr, const double *fl,
                                    const double *bl, int64_t m, const double *strengths,
                                    const double *r_c0s, const double *ages, double nu,
                                    int32_t leg_mask, int32_t precision, double *out);
/* Expanded: out[i][m] = sum_legs v(points[i]; leg of vortex m)       (n, m, 3)
 * Replaces warp_kernels.py:305-330 (ring) and :360-382 (horseshoe). */
int fv_vortex_expanded_host(fv_ctx *ctx, const double *points, int64_t n, const double *br,
                            const double *fr, const double *fl, const double *bl, int64_t m,
                            const double *strengths, const double *r_c0s, const double *ages,
                            double nu, int32_t leg_mask, double *out);
int fv_vortex_expanded_dev(fv_ctx *ctx, const double *points, int64_t n, const double *br,
                           const double *fr, const double *fl, const double *bl, int64_t m,
                           const double *strengths, const double *r_c0s, const double *ages,
                           double nu, int32_t leg_mask, double *out);

/* ------------------------------------------------------- AIC assembly (a6) */
/* aic[i*p + j] = (sum over the 4 legs of ring j at unit strength, zero age)
 *                . normals[i]                                           (p, p)
 * Replaces PteraSoftware's _calculate_wing_wing_influences (expanded with
 * unit strengths, then einsum with unit normals;
 * idea-stage/architecture_analysis.md:12-13); batched analogue
 * warp_fsi/kernels_uvlm.py:21-53. */
int fv_aic_host(fv_ctx *ctx, const double *colloc, const double *normals, int64_t p,
                const double *br, const double *fr, const double *fl, const double *bl,
                const double *r_c0s, double nu, double *aic);
int fv_aic_dev(fv_ctx *ctx, const double *colloc, const double *normals, int64_t p,
               const double *br, const double *fr, const double *fl, const double *bl,
               const double *r_c0s, double nu, double *aic);

/* ------------------------------------------------- dense solve (a7) */
/* Solves A_b x_b = rhs_b for `batch` independent row-major n x n systems
 * (LU with partial pivoting; A overwritten by its factors, rhs by x).
 * info[b] = 0 or k+1 when column k has an exactly zero pivot.  One
 * workgroup per system; n <= 1024 (the factorisation stages n x 17 panels
 * in the 160 KB LDS; n <= 141 keeps the whole FP64 matrix there).
 * Replaces np.linalg.solve in standalone_uvlm.py:427 / PteraSoftware's
 * _calculate_vortex_strengths, and warp_fsi/batched_solver.py:20-63. */
int fv_dense_solve_host(fv_ctx *ctx, int64_t batch, int64_t n, double *A, double *rhs,
                        int32_t *info);
int fv_dense_solve_dev(fv_ctx *ctx, int64_t batch, int64_t n, double *A, double *rhs,
                       int32_t *info);
/* Mixed precision: factorise in FP32, refine with FP64 residuals until the
 * relative correction is below tol or max_iter is reached.  A is not
 * modified.  iters (nullable) receives the refinement steps used per system. */
int fv_dense_solve_mixed_dev(fv_ctx *ctx, int64_t batch, int64_t n, const double *A,
                             double *rhs, int32_t max_iter, double tol, int32_t *info,
                             int32_t *iters);
/* Host twin of fv_dense_solve_mixed_dev (copies in and out, synchronous);
 * info and iters (nullable) have batch entries. */
int fv_dense_solve_mixed_host(fv_ctx *ctx, int64_t batch, int64_t n, const double *A,
                              double *rhs, int32_t max_iter, double tol, int32_t *info,
                              int32_t *iters);

/* ------------------------------------------- particle Biot-Savart (a8, a9) */
/* u(x_t) = -(1/4pi) sum_j g(r/sigma_j)/r^3 (x_t - x_j) x alpha_j        (t, 3)
 * Replaces kernel.py:69-116 / warp_vpm.py:52-73.  The self pair (r < 1e-12)
 * is excluded (warp_vpm.py:37-38). */
int fv_particle_velocity_host(fv_ctx *ctx, const double *targets, int64_t t,
                              const double *pos, const double *alpha, const double *sigma,
                              int64_t n, int32_t kernel, int32_t precision, double *out);
/* J[t][i][j] = dU_i/dx_j                                                  (t, 3, 3)
 * Replaces kernel.py:119-190 / warp_vpm.py:144-173 (self pair excluded). */
int fv_particle_jacobian_host(fv_ctx *ctx, const double *targets, int64_t t,
                              const double *pos, const double *alpha, const double *sigma,
                              int64_t n, int32_t kernel, int32_t precision, double *J);

/* ------------------------------------------- vortex particle field (a10) */
/* Device-resident SoA particle field.  Replaces particles.py:15-246
 * (VortexParticleField(max_particles, nu, rlxf, kernel)). */
int fv_vpm_create(fv_ctx *ctx, int64_t max_particles, double nu, double rlxf, int32_t kernel,
                  int32_t precision, fv_vpm **out);
void fv_vpm_destroy(fv_vpm *f);
/* particles.py:47-57; FV_E_CAP when np + n > max_particles (:50-51). */
int fv_vpm_add_host(fv_vpm *f, const double *pos, const double *gamma, const double *sigma,
                    int64_t n);
int fv_vpm_count(fv_vpm *f, int64_t *np);
/* Copies the first np particles (AoS) to host; any pointer may be NULL. */
int fv_vpm_read_host(fv_vpm *f, double *pos, double *gamma, double *sigma, double *age);
/* Overwrites the first np particles from host arrays (for checkpoint restore). */
int fv_vpm_write_host(fv_vpm *f, int64_t np, const double *pos, const double *gamma,
                      const double *sigma, const double *age);
/* particles.py:72-82 */
int fv_vpm_induce_host(fv_vpm *f, const double *targets, int64_t t, double *out);
/* Rule for particles closer than 1e-12 (FV_PAIRS_*, default SKIP_CLOSE).
 * Velocities and stretching are the same under both up to round-off (their
 * dx = 0 terms vanish); the relaxation's Gaussian Jacobian is not: under
 * FV_PAIRS_CLAMP the self pair adds -2 aux2 alpha to omega with aux2 =
 * -g(1e-12/sigma)/(4 pi 1e-36), g the round-off of erf(rho/sqrt2) -
 * sqrt(2/pi) rho, evaluated bit for bit as the reference's CPU path
 * (particles.py:219 -> kernel.py:156-172). */
int fv_vpm_set_pair_rule(fv_vpm *f, int32_t rule);
/* particles.py:85-199 followed by _relax_pedrizzetti (:202-246), with a
 * constant freestream uinf[3]. */
int fv_vpm_advect_rk3(fv_vpm *f, double dt, const double *uinf, int32_t stretch,
                      int32_t free_wake);
/* The stage velocity of advect_rk3's callables U_inf_func and
 * bound_velocity_func (particles.py:126-133): called at every RK stage with
 * the stage positions (n, 3, host memory); fills u_bound and u_inf (n, 3)
 * and returns 0, or nonzero to abort the step. */
typedef int (*fv_stage_velocity_fn)(void *user, int64_t n, const double *pos, double *u_bound,
                                    double *u_inf);
/* particles.py:85-199 with the callables (U_total = U_part + U_bound + U_inf
 * at every stage) followed by _relax_pedrizzetti (:202-246).  Every stage
 * reads the positions back and uploads the two fields (synchronous); a
 * nonzero return of fn aborts the step with FV_E_ARG. */
int fv_vpm_advect_rk3_ext(fv_vpm *f, double dt, fv_stage_velocity_fn fn, void *user,
                          int32_t stretch, int32_t free_wake);

/* ----------------------------------------- panel -> particle conversion (a11) */
/* Each ring with |gamma| >= 1e-15 yields its leg particles (front, left,
 * back, right; alpha = gamma (b - a), kept if |alpha|^2 > 1e-30) and its
 * gamma is set to 0.  Appends to the field.  Replaces
 * tests/experiment_hybrid_panel_particle.py:145-204. */
int fv_convert_rings_host(fv_vpm *f, int64_t nr, const double *br, const double *fr,
                          const double *fl, const double *bl, double *gamma, double sigma,
                          int64_t *added);

/* Particle population control (newton_pc/adapters/flap.py:466-563).
 * fv_vpm_drop_oldest: beyond max_particles the oldest particles are removed
 * (flap.py:469-473).  fv_vpm_merge_host: pairwise moment-conserving merge
 * (flap.py:478-563): the oldest np - protect particles (or, with
 * protect_dist > 0, those farther than protect_dist from the centre of the
 * nref <= 64 reference points, host xyz) are hashed into cells of size
 * 0.1 + 0.15 d (d: distance to that centre); consecutive aligned pairs of one
 * cell merge into (a1 + a2, |a|-weighted centroid, max sigma) when the
 * velocity change at the reference points stays below eps * vinf_norm;
 * n_merged (may be NULL: no synchronisation) receives the pairs merged by
 * this call.  Both run on the stream without host synchronisation. */
int fv_vpm_drop_oldest(fv_vpm *f, int64_t max_particles);
int fv_vpm_merge_host(fv_vpm *f, const double *ref, int32_t nref, double eps, double vinf_norm,
                      int64_t protect, double protect_dist, int64_t *n_merged);
/* Cumulative merged pairs and candidates skipped for |cell index| >= 2^20. */
int fv_vpm_merge_stats(fv_vpm *f, int64_t *merged, int64_t *out_of_range);
/* Total vector circulation sum(alpha) of the particles (flap.py:565-578). */
int fv_vpm_ledger_host(fv_vpm *f, double *out3);

/* -------------------------------------------------- hybrid step (a12-a14) */
/* One lattice per lifting surface; grids hold the panel corner vertices of
 * every surface for every time step:
 *   grids[((step * total_vertices) + surface_offset + (i * (ns+1) + j)) * 3 + c],
 * i = 0..nc chordwise (leading edge first), j = 0..ns spanwise (left to right). */
typedef struct fv_hybrid_desc {
    int32_t n_surfaces;
    const int32_t *nc;           /* [n_surfaces] chordwise panels */
    const int32_t *ns;           /* [n_surfaces] spanwise panels */
    int64_t num_steps;
    const double *grids;         /* host array, see above */
    double vinf[3];              /* freestream velocity in the geometry frame */
    double rho;
    double dt;
    double nu;                   /* viscosity in the ring core growth law */
    double rc0_bound;            /* core radius of bound rings */
    double rc0_wake;             /* core radius of wake rings */
    int32_t n_keep;              /* wake rows kept as rings (HybridSolver n_keep) */
    double sigma;                /* core size of converted particles */
    int32_t kernel;              /* FV_KERNEL_* for the particle field */
    int32_t free_wake;           /* particles feel each other */
    int32_t stretch;             /* rVPM stretching */
    double rlxf;                 /* Pedrizzetti relaxation factor */
    double vpm_nu;               /* particle viscosity */
    double sref;                 /* reference area for CL / CDi */
    int64_t max_particles;
    int32_t precision;           /* FV_PREC_* for pair sums and the dense solve */
} fv_hybrid_desc;

/* Replaces tests/experiment_hybrid_panel_particle.py:87-143 (HybridSolver over
 * PteraSoftware's UnsteadyRingVortexLatticeMethodSolver step loop). */
int fv_hybrid_create(fv_ctx *ctx, const fv_hybrid_desc *desc, fv_hybrid **out);
void fv_hybrid_destroy(fv_hybrid *h);
/* Advances nsteps time steps on the device (no host synchronisation). */
int fv_hybrid_step(fv_hybrid *h, int64_t nsteps);
int fv_hybrid_current_step(fv_hybrid *h, int64_t *step);
/* Copies CL / CDi histories (num_steps entries; steps not yet run are 0) and
 * the latest bound circulations (total panels).  Any pointer may be NULL. */
int fv_hybrid_read(fv_hybrid *h, double *cl, double *cdi, double *gamma);
/* The particle field driven by the hybrid case (owned by it). */
fv_vpm *fv_hybrid_particles(fv_hybrid *h);
/* Pair counts (segment pairs, particle pairs) evaluated since creation; the
 * particle pairs are counted on the device from the exact particle count of
 * each pass (this call synchronises the context stream). */
int fv_hybrid_pair_counts(fv_hybrid *h, double *seg_pairs, double *particle_pairs);
/* Switch the particle-field flags / the arithmetic for the following steps
 * (HybridSolver(free_vpm=...) and advect_rk3(stretch=...),
 * experiment_hybrid_panel_particle.py:107,140-143). */
int fv_hybrid_set_flags(fv_hybrid *h, int32_t free_wake, int32_t stretch);
/* Population control after every particle advect (newton_pc/adapters/
 * flap.py:466-475): scheme 0 none, 1 keep the newest max_particles, 2 merge
 * (fv_vpm_merge_host with the step's collocation points 0, s, 2s, ... (s =
 * max(1, P / 9), at most 9, flap.py:426) as reference points and |V_inf|). */
int fv_hybrid_set_population(fv_hybrid *h, int32_t scheme, int64_t max_particles, double eps,
                             int64_t protect, double protect_dist);
/* Total vector circulation: particle alpha + stored wake rings' G (b - a)
 * over their legs (flap.py:565-578). */
int fv_hybrid_ledger_host(fv_hybrid *h, double *out3);
/* Ring-wake motion.  free_wake 0: prescribed (every vertex moves by V dt, the
 * default); 1: free (PteraSoftware's prescribed_wake=False path): each wake
 * vertex moves by (V + u) dt with u induced at its pre-step position by the
 * bound rings (this step's circulation) and the wake rings, all vertices from
 * one snapshot.  Rings whose corner-mean x exceeds truncate_x lose their
 * circulation after shedding (warp_fsi/kernels_wake.py:110-126; INFINITY:
 * never). */
int fv_hybrid_set_ring_wake(fv_hybrid *h, int32_t free_wake, double truncate_x);
int fv_hybrid_set_precision(fv_hybrid *h, int32_t precision);
/* Factor reuse of the per-step circulation solve (no counterpart in the
 * reference, which factors every step: standalone_uvlm.py:427): on (the
 * default), a step whose AIC equals the last factored one bit for bit (a
 * lattice that does not move) skips the factorization and runs the two
 * substitutions on the stored factors; the circulation is the same bit for
 * bit.  0 turns it off (every step factors). */
int fv_hybrid_set_solve_reuse(fv_hybrid *h, int32_t on);
/* Steps so far whose circulation solve took the stored factors. */
int fv_hybrid_solve_reuse_count(fv_hybrid *h, int64_t *reused);

/* Graph mode (no counterpart in the reference, whose step is a Python loop,
 * tests/experiment_hybrid_panel_particle.py:125-143; for callers that want one
 * launch per time step).  on: every launch of a step is sized by capacities
 * (the particle field's max_particles) and the kernels read the particle
 * count and the step index from device memory, so a steady step -- ring wake
 * at n_keep + 1 rows, a conversion every step, particles present, not the
 * last step, no closure and no population control -- enqueues the same work
 * as the one before; fv_hybrid_step then runs one steady step eagerly (which
 * sizes every buffer), captures the next one of each wake parity into a
 * hipGraph and replays it (one graph launch per step).  Other steps run
 * eagerly.  A change of flags, precision, ring wake, reuse or the particle
 * field's settings makes the graphs stale; they are captured again.  Graph
 * mode also sets the context's pair-sum plans to their capacity-sized form
 * (fixed shapes, source slices sized on the device), which moves particle
 * sums by round-off against the default plans: graph replays equal eager
 * graph-mode steps bit for bit (on = 2: graph-mode launches, every step
 * eager; 0: off).  Use a context of its own per run. */
int fv_hybrid_set_graph(fv_hybrid *h, int32_t on);
/* Steps fv_hybrid_step replayed from graphs / ran eagerly in graph mode. */
int fv_hybrid_graph_stats(fv_hybrid *h, int64_t *graph_launches, int64_t *eager_steps);

/* Batched runs (a reduced-frequency or configuration sweep; no counterpart in
 * the reference, which runs its cases one after another,
 * tests/experiment_hybrid_panel_particle.py:206-329): n independent runs with
 * uneven panel, wake and particle counts advance together.  Steps where every
 * run is steady and warm (above) replay ONE graph holding a step of every
 * run (each run's work on its own context stream inside the graph), one graph
 * launch per time step for the whole batch; other steps run each run's step
 * eagerly on its own stream.  The runs keep their own state and readers
 * (fv_hybrid_read ...); destroy the batch before its runs. */
int fv_hybrid_batch_create(fv_hybrid *const *runs, int32_t n, fv_hybrid_batch **out);
int fv_hybrid_batch_step(fv_hybrid_batch *b, int64_t nsteps);
int fv_hybrid_batch_stats(fv_hybrid_batch *b, int64_t *graph_launches, int64_t *eager_steps);
void fv_hybrid_batch_destroy(fv_hybrid_batch *b);

/* ------------------------------------------ learned far-field closure MLP */
/* 32 -> 128 -> 128 -> 128 -> 16 SiLU network, BF16 weights / activations on
 * MFMA with FP32 accumulation; FP32 master parameters in the layout
 * W1[128][32] b1[128] W2[128][128] b2[128] W3[128][128] b3[128] W4[16][128]
 * b4[16] (fv_mlp_param_count() floats).  No reference counterpart: the
 * reference has no learned closure; this is the network of the far-field
 * closure that replaces beyond-cutoff particle sums (the exact direct sum of
 * kernel.py:69-190 stays the fallback and the training target). */
typedef struct fv_mlp fv_mlp;
int fv_mlp_create(fv_ctx *ctx, fv_mlp **out);
void fv_mlp_destroy(fv_mlp *m);
int fv_mlp_param_count(void);
int fv_mlp_set_params(fv_mlp *m, const float *params);   /* host; resets Adam */
int fv_mlp_get_params(fv_mlp *m, float *params);         /* host */
/* X (nb x 32), Y (nb x 16) fp32 device arrays; T of the gradient calls is nb x 16. */
int fv_mlp_forward_dev(fv_mlp *m, const float *X, int64_t nb, float *Y);
/* Gradient of L = (1/nb) sum_s sum_j wout_j (Y_sj - T_sj)^2; grad (device,
 * nullable) receives a copy. */
int fv_mlp_gradients_dev(fv_mlp *m, const float *X, const float *T, const float *wout,
                         int64_t nb, float *grad);
/* Adam (beta 0.9 / 0.999, eps 1e-8, bias-corrected) with grad (device; NULL
 * = the last computed gradient). */
int fv_mlp_adam_dev(fv_mlp *m, const float *grad, float lr);
int fv_mlp_train_step_dev(fv_mlp *m, const float *X, const float *T, const float *wout,
                          int64_t nb, float lr);
int fv_mlp_forward_host(fv_mlp *m, const float *X, int64_t nb, float *Y);
int fv_mlp_gradients_host(fv_mlp *m, const float *X, const float *T, const float *wout,
                          int64_t nb, float *grad);
int fv_mlp_adam_host(fv_mlp *m, const float *grad, float lr);

/* Far-field closure of a particle field's self sums (velocity, stretching,
 * vorticity).  Cells of edge `cell`; pairs of cells whose centres are within
 * kappa * cell are summed exactly, farther cells through their first-order
 * expansion (sum alpha, sum alpha (x - xbar)^T) plus the residual predicted by
 * `mlp` (NULL: expansion only).  Every advect trains the network on
 * `train_samples` exact far (particle, cell) sums (`train_iters` Adam steps at
 * `lr`), and every `check_every` advects compares the closure velocity with
 * the exact direct sum: above `tol` (normwise relative) the closure is
 * rejected and the direct sums are used until the next check passes.
 * No reference counterpart (the reference sums directly, kernel.py:69-190). */
int fv_vpm_set_closure(fv_vpm *f, fv_mlp *mlp, double cell, double kappa, int32_t check_every,
                       double tol, int32_t train_samples, int32_t train_iters, double lr);
/* Cells farther than kappa_local * cell from a particle's cell (centre to
 * centre) enter through a second-order local expansion at that cell's centre,
 * built once per binning from every such cell's monopole + dipole, instead of
 * per particle (0: off, the default; else >= 2 kappa, beyond the learned
 * shell).  No counterpart in the reference (the closure is this build's). */
int fv_vpm_set_closure_local(fv_vpm *f, double kappa_local);
int fv_vpm_clear_closure(fv_vpm *f);
int fv_vpm_closure_stats(fv_vpm *f, double *last_err, int64_t *checks, int64_t *rejections,
                         int32_t *active, double *far_evals);
/* Near-field pairs the closure has evaluated so far (every pass: each target
 * particle against the particles of its cell's stencil columns; exact). */
int fv_vpm_closure_near_pairs(fv_vpm *f, double *pairs);
/* Closure velocity at the particles (host, n x 3; nullable) with or without
 * the learned residual and its normwise error against the exact sum. */
int fv_vpm_closure_velocity_host(fv_vpm *f, int32_t use_mlp, double *out, double *err);
/* One closure pass at the particles (original order, n x 3, host): mode 0
 * velocity, 2 stretching J alpha, 3 vorticity; kernel selects the near-field
 * smoothing (the reference's relaxation always uses FV_KERNEL_GAUSSIAN_ERF,
 * particles.py:219).  No reference counterpart (tests of the closure). */
int fv_vpm_closure_sum_host(fv_vpm *f, int32_t mode, int32_t kernel, int32_t use_mlp,
                            double *out);
/* Runs `rounds` online-training rounds on the current particle state. */
int fv_vpm_closure_train(fv_vpm *f, int32_t rounds);

/* ------------------------------------------- batched aeroelastic sweep (f2) */
/* Many independent aeroelastic hybrid UVLM instances (a flutter-speed sweep)
 * stepped together: per step one batched particle-induction launch and one
 * launch with a workgroup per instance for the rest of the step, captured once
 * in a hipGraph and replayed.  Replaces the per-speed loop of
 * tests/benchmark_goland.py:100-243 over AeroelasticSolver
 * (src/fluxvortex/aeroelastic_solver.py:63-221) and BeamFE.step
 * (src/fluxvortex/beam_fe.py:231-277), with the hybrid N_keep wake.
 *
 * Per instance and step s: the lattice = the undeformed grid deformed by the
 * beam's nodal heave w and twist theta (each vertex with w, theta linearly
 * interpolated at its |y|; rotation about the elastic axis x_ea); the hybrid
 * UVLM step (the device driver's statement); for 1 <= s <= num_steps - 2 the
 * panel forces' lift components, summed per spanwise station (station y = the
 * first chordwise panel's collocation y, moment = sum x_colloc * lift about
 * x = 0), go linearly to the beam nodes, one Newmark step advances the beam
 * and its heave / twist deform the lattice of step s + 1 (else step s + 1 is
 * undeformed).  Particles are a prescribed wake convected with V dt.
 * n_keep >= num_steps keeps the whole ring wake (no particles, the reference
 * benchmark's wake).  Every instance holds its ring wake in device memory;
 * a step is two launches (the device-wide pair work of all instances, then
 * the per-instance remainder).
 * Limits: nc <= 16, ns <= 32, nc * ns <= 64, beam nodes <= 17. */
typedef struct fv_sweep_desc {
    int32_t nc, ns;              /* one lattice, (nc+1) x (ns+1) corner grid */
    const double *grid;          /* undeformed corner grid (host), x aft, y span */
    int64_t num_steps;
    double vinf[3];
    double rho, dt;
    double nu;                   /* ring core growth viscosity */
    double rc0_bound, rc0_wake;
    int32_t n_keep;
    double sigma;                /* particle core of converted rings */
    int32_t kernel;              /* FV_KERNEL_* of the particle induction */
    double sref;
    int64_t max_particles;
    int32_t couple;              /* 0: rigid lattice (no structure) */
    int32_t n_nodes;             /* beam nodes incl. the clamped root */
    const double *y_nodes;       /* n_nodes, increasing from the root */
    const double *M, *C, *Keff_inv; /* free-DOF matrices, 3 (n_nodes - 1) square */
    const double *d0, *v0, *a0;  /* free-DOF initial state */
    double beta, gamma;          /* Newmark parameters */
    double x_ea;                 /* elastic axis x (absolute) */
    double relax;                /* deformation relaxation factor (1: none) */
    int32_t precision;           /* FV_PREC_FP64, or FV_PREC_MIXED (FV_PREC_* above) */
    int32_t feedback;            /* 1: the deformed lattice carries the next step's
                                  * aerodynamics; 0: the aerodynamics stay on the
                                  * undeformed lattice (the reference's behaviour) */
    double lift_sign;            /* station lift = lift_sign * F_z (reference: -1) */
    int32_t free_ring_wake;      /* 1: wake vertex rows 0 .. n_keep move with V + the
                                  * induced velocity (frozen snapshot, as
                                  * fv_hybrid_set_ring_wake); 0: prescribed (V dt) */
    int32_t truncate_wake;       /* 1: after shedding, rings whose corner-mean x
                                  * exceeds truncate_x lose their circulation */
    double truncate_x;
} fv_sweep_desc;
typedef struct fv_sweep fv_sweep;
int fv_sweep_create(fv_ctx *ctx, const fv_sweep_desc *descs, int32_t n, int32_t use_graph,
                    fv_sweep **out);
void fv_sweep_destroy(fv_sweep *sw);
/* Advances every unfinished instance by up to `steps` steps. */
int fv_sweep_step(fv_sweep *sw, int64_t steps);
/* Host copies of instance i: cl, cdi, tip_w, tip_theta [num_steps] (tip
 * values are NaN where no coupling ran), gamma [nc*ns], particle count; any
 * pointer may be NULL.  status: 0, or the step + 1 of a zero pivot, or -2 on
 * particle overflow. */
int fv_sweep_read(fv_sweep *sw, int32_t i, double *cl, double *cdi, double *tip_w,
                  double *tip_theta, double *gamma, int64_t *np, int32_t *status);

#ifdef __cplusplus
}
#endif
#endif
