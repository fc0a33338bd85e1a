// Kernel declarations and launch-parameter structs shared by the kernels
// (sph_kernels.hip) and the host orchestration (sph_ctx.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "sph_device.hpp"

namespace sph {

// uniform grid over occupied cells: cell (cx,cy,cz) -> key
// ((cx-org.x)*dim.y + (cy-org.y))*dim.z + (cz-org.z); start has ncell+1 entries
//
// Shared hash buckets: the reference keeps its samples in a HashMap keyed by
// the 64-bit spatial hash of the cell (spatial_hashing.rs:27-103), so two
// cells with equal keys share one bucket, and a stencil holding both visits
// that bucket twice (its samples in range are emitted twice, at each cell's
// place) -- also when one of the two cells is empty or lies outside the
// occupied box. part is null when no cell within two stencil reaches of the
// box shares a key with a cell of the box; otherwise it covers the box
// widened by pm cells on every side, and part[2 e + t] (t = 0, 1) are the box
// keys of the cells of the box that share the bucket of the cell with widened
// key e (kNoCell where there are fewer), and mkey[j - mbase] is the sample
// index (the bucket's fill order) of the member at sorted position j.
constexpr uint32_t kNoCell = 0xffffffffu;
struct GridDev {
    int org[3];
    int dim[3];
    const uint32_t *start;
    const uint32_t *part;
    const uint32_t *mkey;
    uint32_t mbase;
    int pm;
};

// The reference's stencil visit over shared buckets: the (2 reach + 1)^3
// cells in (dx, dy, dz) order, and at each cell every member of its bucket,
// i.e. of the cell and of the cells of the box sharing its key, merged in
// sample-index order. A bucket partner outside the query's stencil holds no
// sample within range (it lies more than a stencil reach away), so only
// partners inside the stencil are merged. fn(j) receives sorted positions as
// stored in start.
template <class Fn>
__device__ void bucket_walk(const GridDev &g, int cx, int cy, int cz, int reach, Fn &&fn) {
    const int ex1 = g.dim[1] + 2 * g.pm, ex2 = g.dim[2] + 2 * g.pm;  // widened box extents
    for (int dx = -reach; dx <= reach; dx++) {
        const int gx = cx + dx - g.org[0];
        if (gx < -g.pm || gx >= g.dim[0] + g.pm) continue;
        for (int dy = -reach; dy <= reach; dy++) {
            const int gy = cy + dy - g.org[1];
            if (gy < -g.pm || gy >= g.dim[1] + g.pm) continue;
            for (int dz = -reach; dz <= reach; dz++) {
                const int gz = cz + dz - g.org[2];
                if (gz < -g.pm || gz >= g.dim[2] + g.pm) continue;
                uint32_t lo[3], hi[3];
                int nl = 0;
                if (gx >= 0 && gx < g.dim[0] && gy >= 0 && gy < g.dim[1] && gz >= 0 &&
                    gz < g.dim[2]) {  // the cell's own members
                    const uint32_t a =
                        ((uint32_t)gx * (uint32_t)g.dim[1] + (uint32_t)gy) * (uint32_t)g.dim[2] +
                        (uint32_t)gz;
                    lo[0] = g.start[a];
                    hi[0] = g.start[a + 1];
                    nl = 1;
                }
                const size_t e = ((size_t)(gx + g.pm) * (size_t)ex1 + (size_t)(gy + g.pm)) * (size_t)ex2 +
                                 (size_t)(gz + g.pm);
                for (int t = 0; t < 2; t++) {
                    const uint32_t b = g.part[2 * e + t];
                    if (b == kNoCell) continue;
                    const int bz = (int)(b % (uint32_t)g.dim[2]) + g.org[2];
                    const int by = (int)((b / (uint32_t)g.dim[2]) % (uint32_t)g.dim[1]) + g.org[1];
                    const int bx = (int)(b / ((uint32_t)g.dim[2] * (uint32_t)g.dim[1])) + g.org[0];
                    if (abs(bx - cx) > reach || abs(by - cy) > reach || abs(bz - cz) > reach) continue;
                    lo[nl] = g.start[b];
                    hi[nl] = g.start[b + 1];
                    nl++;
                }
                for (;;) {  // merge the member lists (each in sample-index order)
                    int best = -1;
                    uint32_t bk = 0;
                    for (int t = 0; t < nl; t++)
                        if (lo[t] < hi[t]) {
                            const uint32_t k = g.mkey[lo[t] - g.mbase];
                            if (best < 0 || k < bk) {
                                best = t;
                                bk = k;
                            }
                        }
                    if (best < 0) break;
                    fn(lo[best]);
                    lo[best]++;
                }
            }
        }
    }
}

struct SolveConst {
    double dt;
    double V0;
    double eps;    // min_diagonal_element
    double omega;  // relaxation_factor
    double gamma;  // boundary_pressure_acceleration_weighting
};

__global__ void k_fill_u32(uint32_t *p, size_t n, uint32_t v);
__global__ void k_fill_f64(double *p, size_t n, double v);
__global__ void k_cells(size_t n, double cs, const double *x, const double *y, const double *z,
                        int *ci, int *bbox_partials);
__global__ void k_cells_aos(size_t n, double cs, const double *p, int *ci, int *bbox_partials);
__global__ void k_fold_bbox(const int *partials, int nparts, int *bbox);
__global__ void k_init_bbox(int *bbox);
__global__ void k_hash_check(GridDev g, int reach, unsigned int *flag);
__global__ void k_hash_partners(GridDev g, int span, uint32_t *part, unsigned int *flags);
__global__ void k_cell_count(size_t n, const int *ci, GridDev g, uint32_t *count);
__global__ void k_cell_scatter(size_t n, const int *ci, GridDev g, const uint32_t *start,
                               uint32_t *cursor, uint32_t *perm);
__global__ void k_cell_order(size_t ncell, const uint32_t *start, uint32_t *perm,
                             const uint32_t *ref);
template <class T>
__global__ void k_gather(size_t n, const uint32_t *perm, const T *src, T *dst);
template <class T>
__global__ void k_scatter(size_t n, const uint32_t *perm, const T *src, T *dst);
__global__ void k_gather_aos(size_t n, const uint32_t *perm, const double *p, double *x,
                             double *y, double *z);
template <bool FILL>
__global__ void k_neighbors_aos(size_t nq, const double *q, double cs, int reach, double dthr,
                                GridDev g, const double *sx, const double *sy, const double *sz,
                                const uint32_t *off, uint32_t *out_count, const uint32_t *perm,
                                uint64_t *out_idx);
template <bool FILL>
__global__ void k_neighbors(size_t nq, const double *qx, const double *qy, const double *qz,
                            double cs, int reach, double dthr, GridDev g, const double *sx,
                            const double *sy, const double *sz, const uint32_t *off,
                            uint32_t *out_count, uint32_t *out_idx);
__global__ void k_scan_block(const uint32_t *in, size_t n, uint32_t *out, uint32_t *block_sums);
__global__ void k_scan_add(uint32_t *out, size_t n, const uint32_t *block_offsets);
__global__ void k_static_volume(size_t n, KernelConst K, double gamma_v, const double *x,
                                const double *y, const double *z, const uint32_t *off,
                                const uint32_t *idx, double *vol);
__global__ void k_map_indices(size_t n, const uint32_t *perm, const uint32_t *in, uint64_t *out);

// volume-map construction
// Sensor queries (SPHSystem::get_quantity_at_positions, sph/mod.rs:241-341):
// the fluid in (cell, reference index) order through `perm`, read in place
struct QueryFluid {
    const double *x, *y, *z, *vx, *vy, *vz, *m, *V, *p;
    const uint32_t *perm;  // sorted slot -> storage index
    GridDev g;
};
struct QueryConst {
    KernelConst K;
    double V0, cs, dthr, dxr, vmax;
    int reach;
};
enum { kQSpeed = 0, kQVolume, kQDensity, kQDensityError, kQPressure, kQKineticEnergy };
template <int Q>
__global__ void k_quantity(size_t nq, const double *qx, const double *qy, const double *qz,
                           QueryConst C, QueryFluid F, int nb, const BodyDev *bodies,
                           const GridDev *sgrid, float *out);

constexpr int kVmMaxDegree = 32;
struct VmGrid {
    double x_min[3];
    int n[3];  // cells per axis of the volume-map lattice
    double dx;
};
struct VmQuad {
    double h, bv, wfac;  // support radius, ball volume, alpha / h^3
    int degree;
    const double *w;               // Gauss-Legendre weights on [-1, 1]
    const double *r;               // radial nodes on [0, h]
    const double *st, *ct;         // sin, cos of the polar nodes on [0, pi]
    const double *sp, *cp;         // sin, cos of the azimuthal nodes on [0, 2 pi]
};
__global__ void k_vm_classify(LatticeDev sdf, VmGrid g, double h, double bv, double *values,
                              uint8_t *status, uint32_t *band, unsigned int *nband);
__global__ void k_vm_quadrature(LatticeDev sdf, VmGrid g, VmQuad Q, const uint32_t *band,
                                double *values, uint8_t *status);
__global__ void k_vm_cells(VmGrid g, const uint8_t *status, uint8_t *present);
__global__ void k_vm_finish(size_t n, const uint8_t *present, double *values);
__global__ void k_scan_total(const uint32_t *in, size_t n, uint32_t *out);
__global__ void k_widen_u32(size_t n, const uint32_t *in, uint64_t *out);

// Signed distance to a triangle mesh at lattice nodes
// (TriangleMeshWrapper::signed_distance, volume_map_boundary.rs:1655-1671)
constexpr int kMeshTile = 128;  // triangles staged in LDS per round
struct MeshDev {
    const double *tri;     // [nt][9] vertex positions a, b, c, in tile order
    const uint32_t *orig;  // [nt] mesh index of each triangle in tile order
    const double *box;     // [tiles][6] bounding box (min, max) of each tile's vertices
    const uint32_t *tv;    // [nt][3] vertex ids (mesh order)
    const uint32_t *te;    // [nt][3] edge ids of (a b), (b c), (c a) (mesh order)
    const double *fn;      // [nt][3] unit face normals (mesh order)
    const double *vn;      // [nv][3] vertex pseudo-normals
    const double *en;      // [ne][3] edge pseudo-normals
    uint32_t nt;
    double slack;          // absolute slack of the tile-skip test (rounding of distances)
};
__global__ void k_mesh_sdf(MeshDev M, VmGrid g, double *values);
__global__ void k_sdf_cells(VmGrid g, const double *values, double prune, uint8_t *present);

// Rigid-rigid contact cells (VolumeMapBoundary::generate_contact_cells,
// volume_map_boundary.rs:98-161, 702-768, 835-882): candidate sub-cells of
// the overlapping world AABBs of a body pair, numbered pair by pair in
// (z, y, x) order from `first`
constexpr int kContactMaxDegree = 16;
constexpr int kContactBlock = 256;
// per active cell: volume, sum of weighted positions[3], f_a[3], tau_a[3],
// f_b[3], tau_b[3], quadrature points kept
constexpr int kContactSums = 17;
struct ContactPairDev {
    int a, b;
    uint32_t first;
    uint32_t nn[3];
    double omn[3], cs[3];
};
struct ContactQuadDev {
    KernelConst K;
    int degree;
    double x[kContactMaxDegree], w[kContactMaxDegree];  // Gauss-Legendre rule on [-1, 1]
};
__global__ void k_contact_cells(const ContactPairDev *pairs, int npairs, const BodyDev *bodies,
                                ContactQuadDev Q, uint32_t base, unsigned int *nact,
                                uint32_t cap, uint32_t *cell_idx, double *sums);

}  // namespace sph
