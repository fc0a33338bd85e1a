// Kernel declarations and launch-parameter structs shared by the kernels
// (sph_kernels.hip) and the host orchestration (sph_ctx.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "sph_device.hpp"

namespace sph {

// uniform grid over occupied cells: cell (cx,cy,cz) -> key
// ((cx-org.x)*dim.y + (cy-org.y))*dim.z + (cz-org.z); start has ncell+1 entries
struct GridDev {
    int org[3];
    int dim[3];
    const uint32_t *start;
};

struct NonPressureConst {
    int gravity_mode;
    double gcenter[3];
    double nu_f;
    double nu_b;
    double eps2;  // 0.01 dx^2
};

struct SolveConst {
    double dt;
    double V0;
    double eps;    // min_diagonal_element
    double omega;  // relaxation_factor
    double gamma;  // boundary_pressure_acceleration_weighting
};

__global__ void k_fill_u32(uint32_t *p, size_t n, uint32_t v);
__global__ void k_fill_f64(double *p, size_t n, double v);
__global__ void k_reduce_rows(const double *in, int cols, int rows, double *out, int accumulate);
__global__ void k_integrate_ec(size_t n, double dt, double *x, double *y, double *z, double *vx,
                               double *vy, double *vz, const double *ax, const double *ay,
                               const double *az);
__global__ void k_cells(size_t n, double cs, const double *x, const double *y, const double *z,
                        int *ci, int *bbox_partials);
__global__ void k_fold_bbox(const int *partials, int nparts, int *bbox);
__global__ void k_init_bbox(int *bbox);
__global__ void k_speed_below(size_t n, const double *vx, const double *vy, const double *vz,
                              const double *z, double thr, double *partials);
__global__ void k_fold_speed(const double *partials, int nparts, double *out);
__global__ void k_hash_check(GridDev g, int reach, unsigned int *flag);
__global__ void k_cell_count(size_t n, const int *ci, GridDev g, uint32_t *count);
__global__ void k_cell_scatter(size_t n, const int *ci, GridDev g, const uint32_t *start,
                               uint32_t *cursor, uint32_t *perm);
__global__ void k_cell_order(size_t ncell, const uint32_t *start, uint32_t *perm,
                             const uint32_t *ref);
template <class T>
__global__ void k_gather(size_t n, const uint32_t *perm, const T *src, T *dst);
template <class T>
__global__ void k_scatter_back(size_t n, const uint32_t *perm, const T *src, T *dst);
template <bool FILL>
__global__ void k_neighbors(size_t nq, const double *qx, const double *qy, const double *qz,
                            double cs, int reach, double dthr, GridDev g, const double *sx,
                            const double *sy, const double *sz, const uint32_t *off,
                            uint32_t *out_count, uint32_t *out_idx);
__global__ void k_scan_block(const uint32_t *in, size_t n, uint32_t *out, uint32_t *block_sums);
__global__ void k_scan_add(uint32_t *out, size_t n, const uint32_t *block_offsets);
__global__ void k_vm_sample(size_t n, const double *x, const double *y, const double *z,
                            BodyDev b, double range, double dxr, double vmax, uint8_t *flag,
                            double *bx, double *by, double *bz, double *bvx, double *bvy,
                            double *bvz, double *bvol);
__global__ void k_count_ell(size_t nq, const double *qx, const double *qy, const double *qz,
                            double cs, int reach, double dthr, GridDev g, uint32_t *count,
                            uint32_t *slice_len);
__global__ void k_fill_ell(size_t n, double cs, int reach, double dthr, GridDev g,
                           const double *x, const double *y, const double *z,
                           const uint32_t *slice_off, uint32_t *nj);
__global__ void k_pairs_volume(size_t n, KernelConst K, double V0, const double *x,
                               const double *y, const double *z, const uint32_t *cnt,
                               const uint32_t *slice_off, const uint32_t *nj, double *gxo,
                               double *gyo, double *gzo, const BodyDev *bodies, int nb,
                               double *V);
__global__ void k_static_volume(size_t n, KernelConst K, double gamma_v, const double *x,
                                const double *y, const double *z, const uint32_t *off,
                                const uint32_t *idx, double *vol);
__global__ void k_nonpressure(size_t n, KernelConst K, NonPressureConst P, const double *x,
                              const double *y, const double *z, const double *vx,
                              const double *vy, const double *vz, const double *m,
                              const double *V, const uint32_t *cnt, const uint32_t *slice_off,
                              const uint32_t *nj, const double *gxa, const double *gya,
                              const double *gza, const BodyDev *bodies, int nb, double *ax,
                              double *ay, double *az, double *body_partials);
__global__ void k_sesph_pressure(size_t n, double k, double V0, const double *V, double *p);
__global__ void k_pressure_accel(size_t n, KernelConst K, double gamma, const double *x,
                                 const double *y, const double *z, const double *m,
                                 const double *V, const double *p, const uint32_t *cnt,
                                 const uint32_t *slice_off, const uint32_t *nj,
                                 const double *gxa, const double *gya, const double *gza,
                                 const BodyDev *bodies, int nb, int overwrite, int mirror,
                                 double *tx_, double *ty_, double *tz_, double *body_partials);
__global__ void k_pred_vel(size_t n, double dt, int to_pred, const double *vx, const double *vy,
                           const double *vz, const double *ax, const double *ay,
                           const double *az, double *svx, double *svy, double *svz);
__global__ void k_source_diag(size_t n, KernelConst K, SolveConst S, int vp, const double *x,
                              const double *y, const double *z, const double *svx,
                              const double *svy, const double *svz, const double *m,
                              const double *V, const uint32_t *cnt, const uint32_t *slice_off,
                              const uint32_t *nj, const double *gxa, const double *gya,
                              const double *gza, const BodyDev *bodies, int nb, double *s_f,
                              double *a_ff, double *p);
__global__ void k_jacobi(size_t n, KernelConst K, SolveConst S, const double *x, const double *y,
                         const double *z, const double *V, const double *pax, const double *pay,
                         const double *paz, const double *s_f, const double *a_ff,
                         const uint32_t *cnt, const uint32_t *slice_off, const uint32_t *nj,
                         const double *gxa, const double *gya, const double *gza,
                         const BodyDev *bodies, int nb, double *p, double *err_partials);
__global__ void k_pred_pos(size_t n, double dt, const double *x, const double *y,
                           const double *z, const double *svx, const double *svy,
                           const double *svz, double *sx, double *sy, double *sz);
__global__ void k_ost_commit(size_t n, double dt, const double *ax, const double *ay,
                             const double *az, double *sx, double *sy, double *sz, double *svx,
                             double *svy, double *svz);
__global__ void k_measure(size_t n, double V0, const double *V, const double *m, const double *vx,
                          const double *vy, const double *vz, double *partials);
__global__ void k_map_indices(size_t n, const uint32_t *perm, const uint32_t *in, uint64_t *out);
__global__ void k_widen_u32(size_t n, const uint32_t *in, uint64_t *out);

}  // namespace sph
