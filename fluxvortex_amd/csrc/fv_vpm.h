// fv_vpm.h -- internal layout of the device-resident vortex particle field.
#pragma once

#include "fv_launch.h"

struct fv_closure;
struct fv_mlp;

// SoA particle state (one allocation), RK3 low-storage registers and the
// pair-sum buffer.  np_dev is the authoritative particle count; the host keeps
// an exact copy when it changed the count itself, else an upper bound.
struct fv_vpm {
    fv_ctx *ctx = nullptr;
    int64_t cap = 0;
    double nu = 0.0, rlxf = 0.3;
    int kernel = FV_KERNEL_GAUSSIAN_ERF;
    int precision = FV_PREC_FP64;
    int pair_rule = 0;             // FV_PAIRS_SKIP_CLOSE or FV_PAIRS_CLAMP (relaxation)
    void *block = nullptr;
    double *x = nullptr, *y = nullptr, *z = nullptr;
    double *gx = nullptr, *gy = nullptr, *gz = nullptr;
    double *sig = nullptr, *age = nullptr;
    double *rkX = nullptr, *rkG = nullptr, *rks = nullptr;   // 3cap, 3cap, cap
    double *sum = nullptr;                                   // 9cap pair sums
    double *stage = nullptr;                                 // 9cap host-API staging
    int64_t *np_dev = nullptr;
    int32_t *overflow_dev = nullptr;
    int64_t np_host = 0;
    bool np_exact = true;
    int64_t np_upper = 0;
    double particle_pairs = 0.0;   // pairs of host-counted passes (closure checks)
    double *pairs_dev = nullptr;   // [0] direct self-sum pairs, [1] induced-target pairs,
                                   // counted on the device from the exact particle count
    int64_t state_epoch = 0;       // bumped by every enqueued change of the particle state
    // launches sized by the capacity instead of np_upper (the device count
    // bounds the work): every launch is the same from step to step
    bool bound_cap = false;
    fv_closure *closure = nullptr; // far-field closure (optional)
    void *pop = nullptr;           // population-control scratch (fv_population.hip), lazily made
};

namespace fv {

// The particle count that sizes launches (np_upper, or the capacity with
// bound_cap; 0 while the field is empty).
inline int64_t launch_count(const fv_vpm *f)
{
    return f->np_upper <= 0 ? 0 : (f->bound_cap ? f->cap : f->np_upper);
}

// Enqueues one advect_rk3 + relaxation step (particles.py:85-246).
// With ext (nullable) the freestream and the bound-vortex velocity come from
// the host callback at every stage (uinf unused).
int vpm_advect(fv_vpm *f, double dt, const double *uinf, int stretch, int free_wake,
               fv_stage_velocity_fn ext = nullptr, void *user = nullptr);

// Enqueues the conversion of nr rings (AoS corners on the device) into
// particles appended to the field; converted rings get gamma = 0.
int vpm_convert_rings(fv_vpm *f, int64_t nr, const double *br, const double *fr,
                      const double *fl, const double *bl, double *gamma, double sigma);

// Enqueues u(targets) from the field into out (n x 3); targets AoS.
int vpm_induce(fv_vpm *f, const double *targets, int64_t n, int precision, double *out,
               int pslot);

// Makes np_host exact (synchronises).
int vpm_sync_count(fv_vpm *f);

// Population control and the circulation ledger (fv_population.hip).
int vpm_drop_oldest(fv_vpm *f, int64_t max_particles);
int vpm_merge(fv_vpm *f, const double *ref_dev, int nref, double eps, double vinf_norm,
              int64_t protect, double protect_dist);
int vpm_merge_stats(fv_vpm *f, int64_t *merged, int64_t *out_of_range);
// Merge with the hybrid driver's reference points (9 strided collocation
// points, flap.py:426).
int vpm_merge_colloc(fv_vpm *f, const double *colloc, int64_t P, double eps, double vinf_norm,
                     int64_t protect, double protect_dist);
int vpm_ledger(fv_vpm *f, int64_t nr, const double *br, const double *fr, const double *fl,
               const double *bl, const double *g, double *out3);
void pop_destroy(fv_vpm *f);

// Far-field closure hooks (fv_closure.hip).
// Per-step bookkeeping at the start of an advect: online training and the
// periodic check against the exact sum (sets whether the closure is used).
int closure_step(fv_vpm *f);
// True when the closure replaces the direct self sums for this step.
bool closure_active(const fv_vpm *f);
// Self sum at the particles (original order): mode 0 velocity, 2 stretching,
// 3 vorticity; kernel selects the near-field smoothing.
int closure_self_sum(fv_vpm *f, int mode, int kernel, double *out);
void closure_destroy(fv_closure *cl);

}  // namespace fv
