// Multi-scene launches of the per-step passes.
//
// A batch holds one or more independent systems ("scenes"). Each pass is one
// launch over the concatenation of all scenes' workgroups: workgroup lb maps to
// scene s with prefix[s] <= lb < prefix[s+1] (a scene's workgroups never
// straddle another scene, so per-scene reductions stay deterministic block
// sums). Scenes keep their own particle arrays; the cell table, the pair
// cache and the per-particle sort scratch are batch-global, addressed through
// per-scene bases.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/sph_hip.h"
#include "sph_device.hpp"
#include "sph_kernels.hpp"

namespace sph {

constexpr int kMaxBodiesPerScene = kMaxBodies;

struct SceneView {
    uint32_t n;           // particles
    uint32_t part_base;   // first index of this scene in batch-global per-particle arrays
    uint32_t slice_base;  // first 64-particle slice in the batch-global slice arrays
    uint32_t cell_base;   // first cell in the batch-global cell table
    uint32_t body_base;   // first body in the batch body table
    uint32_t nb;          // bodies of this scene
    int integrator;       // SPH_INTEGRATOR_*
    int reach;
    int gravity_mode;
    KernelConst K;
    SolveConst S;         // dt, V0, min diagonal, omega, gamma
    double V0, cs, dthr, dxr, vmax_ball, disable_below, stiffness;
    double nu_f, nu_b, eps2;
    double target_err;    // target average density error (percent)
    double gcenter[3];
    GridDev grid;         // occupied-cell box; start = batch table + cell_base
    double *x, *y, *z, *vx, *vy, *vz, *ax, *ay, *az, *m, *V, *p;
    double *q4;   // {x, y, z, V} per particle: the record neighbor loops gather
    double *pa3;  // {pressure acceleration} per particle (24 bytes): the Jacobi scratch
    double *sv4;  // {predicted velocity, m} per particle: gathered by the source pass
    double *v3;   // {velocity} per particle (24 bytes): gathered by the viscosity pass
    double *s_f, *a_ff, *svx, *svy, *svz, *sxp, *syp, *szp;
    double *ix, *iy, *iz;  // the integrator's position slot (Verlet: x one step back)
    int keep_self;         // the pair cache lists the particle itself (SESPHwSplitting's density)
    double *tx, *ty, *tz, *tvx, *tvy, *tvz, *tm;  // permutation targets
    uint32_t *fid, *ref, *tfid, *tref;
};

struct BatchDev {
    const SceneView *views;
    const uint32_t *pblk;  // workgroup prefix over particles (ns + 1)
    const uint32_t *cblk;  // workgroup prefix over cells (ns + 1)
    int ns;
    // a launch may cover a subset of the scenes: entry k of pblk then belongs
    // to scene smap[k] (smap == nullptr: all scenes, in order)
    const uint32_t *smap;
    int ns_all;
    int *ci;               // 3 per particle
    uint32_t *perm;
    uint32_t *cnt;
    // wavefront-sliced ELL pair cache: slice s holds 64 consecutive sorted particles of
    // one scene, and the k-th neighbor of lane l sits at slot
    // 64 slice_off[s] + 128 (k / 2) + 2 l + k % 2 (rows stored in pairs), so the
    // accesses of neighbors k and k+1 of a wavefront are one contiguous run (512 B of
    // indices); slice_len[s] = the slice's largest neighbor count rounded up to even
    uint32_t *slice_len;
    const uint32_t *slice_off;
    uint32_t *ell_j;                   // neighbor's scene-local sorted index
    double *ell_gx, *ell_gy, *ell_gz;  // grad W(x_i - x_j)
    uint32_t *cell_cnt;
    uint32_t *cell_start;
    uint32_t *cursor;
    const BodyDev *bodies;
    double *partials;      // per workgroup
    // pressure solve state per scene, advanced on the device: a scene leaves
    // the Jacobi loop on its own termination test
    uint32_t *active;      // still iterating
    uint32_t *iters;       // iterations done
    double *chunk_err;     // [k][scene] predicted density error of iteration k of a chunk
    uint32_t *chunk_ran;   // [k][scene] 1 if the scene ran iteration k of the chunk
};

// Pair passes take CACHED = true to read grad W from the pair cache (written by
// kb_pairs_volume<true>) or false to recompute it from positions, and a code
// shape: the neighbor loop's unroll factor, a floor on waves per SIMD (which
// caps registers per lane) and the loop form (neighbor_loop's PIPE). Every
// shape evaluates the same expressions in the same order, so results do not
// depend on the shape.
#define SPH_PAIR_VARIANTS(X)                                                      \
    X(4, 0, 1) X(2, 0, 1) X(4, 4, 1) X(1, 0, 0) X(2, 0, 0) X(4, 0, 0) X(4, 0, 3)  \
    X(2, 0, 3) X(4, 3, 0)
constexpr int kPairVariants = SPH_SHAPES;
__global__ void kb_integrate_ec(BatchDev B);
__global__ void kb_speed_below(BatchDev B);
__global__ void kb_cells(BatchDev B);
__global__ void kb_cell_count(BatchDev B);
__global__ void kb_cell_scatter(BatchDev B);
__global__ void kb_cell_order(BatchDev B);
__global__ void kb_permute(BatchDev B);
__global__ void kb_count_ell(BatchDev B);
// neighbor rows a wavefront of kb_sweep_ell buffers in LDS (24 KB per
// workgroup, which leaves room for six workgroups per CU)
constexpr int kSweepRows = 24;
// candidates whose positions kb_sweep_ell loads together
constexpr int kSweepBatch = 4;
__global__ void kb_sweep_ell(BatchDev B, uint32_t cap, uint32_t *maxc);
__global__ void kb_fixed_slices(uint32_t *slice_off, size_t n, uint32_t cap);
__global__ void kb_fill_ell(BatchDev B);
__global__ void kb_vm_sample(BatchDev B);
template <bool CACHED, int UNROLL, int WAVES, int PIPE>
__global__ void kb_pairs_volume(BatchDev B);
template <bool CACHED, int UNROLL, int WAVES, int PIPE>
__global__ void kb_nonpressure(BatchDev B);
__global__ void kb_sesph_pressure(BatchDev B);
__global__ void kb_sesph_split_density(BatchDev B);
template <bool CACHED, int UNROLL, int WAVES, int PIPE>
__global__ void kb_pressure_accel(BatchDev B, int target_scratch, int overwrite, int mirror);
__global__ void kb_pred_vel(BatchDev B, int to_pred);
template <bool CACHED, int UNROLL, int WAVES, int PIPE>
__global__ void kb_source_diag(BatchDev B, int vp);
template <bool CACHED, int UNROLL, int WAVES, int PIPE>
__global__ void kb_jacobi(BatchDev B);
// the Jacobi chunk's kernels reading the batch from device memory over a
// grid rounded up from B->pblk[B->ns] workgroups (B->ns for the fold): the
// launches a captured graph replays whatever the batch's layout
template <bool CACHED, int UNROLL, int WAVES, int PIPE>
__global__ void kb_pressure_accel_d(const BatchDev *B, int target_scratch, int overwrite,
                                    int mirror);
template <bool CACHED, int UNROLL, int WAVES, int PIPE>
__global__ void kb_jacobi_d(const BatchDev *B);
__global__ void kb_fold_converge_d(const BatchDev *B, int after, int k);
__global__ void kb_pred_pos(BatchDev B);
__global__ void kb_ost_commit(BatchDev B);
__global__ void kb_measure(BatchDev B);
__global__ void kb_pack_volume(BatchDev B);
// per-scene folds of per-workgroup partials: out[s*rows + r] = sum (or max) over the
// scene's workgroups of partials[lb*rows + r]
__global__ void kb_fold_sum(const double *partials, const uint32_t *blk, int ns, int rows,
                            double *out);
// fold of the Jacobi error partials per scene + the termination test
// (iisph.rs:19-22, 386-393, 530-531): iteration k of the current chunk
__global__ void kb_fold_converge(BatchDev B, int after, int k);
__global__ void kb_fold_speed(const double *partials, const uint32_t *blk, int ns, double *out);
__global__ void kb_fold_bbox(const int *partials, const uint32_t *blk, int ns, int *out);

}  // namespace sph
