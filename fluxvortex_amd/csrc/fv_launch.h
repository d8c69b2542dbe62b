// fv_launch.h -- internal host-side launchers shared by the entry points
// (segment tables, pair sums, fills).  Not part of the C-ABI.
#pragma once

#include "fv_common.h"
#include "fv_pairs.h"

namespace fv {

struct PairPlan {
    int block;      // 64 or 256 lanes per workgroup
    int tpt;        // targets per lane (1, 2 or 4)
    int64_t gx;     // target blocks
    int split;      // source chunks (blockIdx.y)
    int64_t chunk;  // sources per chunk (multiple of block)
};

// The default shape (by problem size) and an explicit one: `per_cu`
// workgroups per CU are aimed for by splitting the sources.
PairPlan plan_pairs(fv_ctx *ctx, int64_t n_tgt, int64_t n_src);
PairPlan plan_pairs_shape(fv_ctx *ctx, int64_t n_tgt, int64_t n_src, int block, int tpt,
                          int per_cu);

// Fills n doubles with v (a kernel, so it is safe under graph capture).
int fill(fv_ctx *ctx, double *p, int64_t n, double v);
__global__ void k_fill_i32(int32_t *__restrict__ p, int64_t n, int32_t v);

// Builds the segment table of m vortices x legs(mask), vortex-major
// (segment m*nl + k is the k-th selected leg of vortex m).  unit_strength
// ignores gamma/age (AIC: strength 1, age 0).
int build_segtab(fv_ctx *ctx, const double *br, const double *fr, const double *fl,
                 const double *bl, int64_t m, const double *gamma, const double *rc0,
                 const double *age, double nu, int mask, bool unit_strength, SegRec *out);

// out[i] (n x 3) = sum over nseg segments at pts[i] (AoS).  Uses the partial
// workspace slot `pslot` when the source range is split.
int seg_sum(fv_ctx *ctx, const double *pts, int64_t n, const SegRec *segs, int64_t nseg,
            int precision, double *out, int pslot);

// Particle pair sums: out (n_tgt x comp) with comp 3 (velocity, stretching,
// vorticity) or 9 (Jacobian).  mode: 0 velocity, 1 Jacobian, 2 stretching,
// 3 vorticity.  dyn_tgt / dyn_src: optional device counts (upper bounds in
// n_tgt / n_src size the launch).  family names the launch for kernel timing
// (nullptr: "particle_vel" / "particle_jac" / "particle_stretch" / "particle_curl").
int part_sum(fv_ctx *ctx, int kernel, int precision, int mode, const PartSrcArgs &src,
             const PartTgtArgs &tgt, int64_t n_tgt, int64_t n_src, const int64_t *dyn_tgt,
             const int64_t *dyn_src, double *out, int pslot, const char *family = nullptr);

// Fused AIC assembly from a unit-strength ring table (4 legs per ring).
int aic_assemble(fv_ctx *ctx, const double *colloc, const double *normals, int64_t p,
                 const SegRec *unit_segs, double *aic);

inline int legs_in_mask(int mask)
{
    return ((mask >> 0) & 1) + ((mask >> 1) & 1) + ((mask >> 2) & 1) + ((mask >> 3) & 1);
}

}  // namespace fv
