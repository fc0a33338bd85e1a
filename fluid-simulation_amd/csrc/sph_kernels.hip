// Kernels of the SPH particle hot path for gfx950 (MI355X).
//
// Layout in HBM: structure-of-arrays, binary64, particles kept sorted by
// (cell key, reference index) so that the 27-cell neighborhood of a cell is
// 9 contiguous index ranges (z is the fastest key axis) and neighbor reads of
// a wavefront hit the same cache lines.
#include <hip/hip_runtime.h>

#include "sph_device.hpp"
#include "sph_kernels.hpp"

namespace sph {

// ------------------------------------------------------------------------
// fills
// ------------------------------------------------------------------------
__global__ void k_fill_u32(uint32_t *p, size_t n, uint32_t v) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_fill_f64(double *p, size_t n, double v) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// ------------------------------------------------------------------------
// cell coordinates and the occupied-cell bounding box of a point set
// ------------------------------------------------------------------------
// per-block partials of the occupied-cell box (min x,y,z; max x,y,z); a
// single-block pass folds them (no contended atomics on six words)
__device__ __forceinline__ void bbox_block_partial(bool valid, int a, int b, int c,
                                                   int *bbox_partials) {
    __shared__ int lds[6][kBlock / 64];
    int v[6] = {2147483647, 2147483647, 2147483647, -2147483647 - 1, -2147483647 - 1,
                -2147483647 - 1};
    if (valid) {
        v[0] = v[3] = a;
        v[1] = v[4] = b;
        v[2] = v[5] = c;
    }
    for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
        for (int k = 0; k < 3; k++) v[k] = min(v[k], __shfl_xor(v[k], off));
#pragma unroll
        for (int k = 3; k < 6; k++) v[k] = max(v[k], __shfl_xor(v[k], off));
    }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0)
        for (int k = 0; k < 6; k++) lds[k][w] = v[k];
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int q = 1; q < kBlock / 64; q++) {
            for (int k = 0; k < 3; k++) v[k] = min(v[k], lds[k][q]);
            for (int k = 3; k < 6; k++) v[k] = max(v[k], lds[k][q]);
        }
        for (int k = 0; k < 6; k++) bbox_partials[(size_t)blockIdx.x * 6 + k] = v[k];
    }
}

__global__ void k_cells(size_t n, double cs, const double *x, const double *y, const double *z,
                        int *ci, int *bbox_partials) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    int a = 0, b = 0, c = 0;
    if (i < n) {
        a = sat_i32(floor(x[i] / cs));
        b = sat_i32(floor(y[i] / cs));
        c = sat_i32(floor(z[i] / cs));
        ci[3 * i] = a;
        ci[3 * i + 1] = b;
        ci[3 * i + 2] = c;
    }
    bbox_block_partial(i < n, a, b, c, bbox_partials);
}

// the same for [n][3] positions; ci may be null (box only)
__global__ void k_cells_aos(size_t n, double cs, const double *p, int *ci, int *bbox_partials) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    int a = 0, b = 0, c = 0;
    if (i < n) {
        a = sat_i32(floor(p[3 * i] / cs));
        b = sat_i32(floor(p[3 * i + 1] / cs));
        c = sat_i32(floor(p[3 * i + 2] / cs));
        if (ci) {
            ci[3 * i] = a;
            ci[3 * i + 1] = b;
            ci[3 * i + 2] = c;
        }
    }
    bbox_block_partial(i < n, a, b, c, bbox_partials);
}

__global__ void k_fold_bbox(const int *partials, int nparts, int *bbox) {
    __shared__ int lds[6][kBlock];
    int v[6] = {2147483647, 2147483647, 2147483647, -2147483647 - 1, -2147483647 - 1,
                -2147483647 - 1};
    for (int q = threadIdx.x; q < nparts; q += kBlock) {
        for (int k = 0; k < 3; k++) v[k] = min(v[k], partials[(size_t)q * 6 + k]);
        for (int k = 3; k < 6; k++) v[k] = max(v[k], partials[(size_t)q * 6 + k]);
    }
    for (int k = 0; k < 6; k++) lds[k][threadIdx.x] = v[k];
    __syncthreads();
    for (int s = kBlock / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            for (int k = 0; k < 3; k++) lds[k][threadIdx.x] = min(lds[k][threadIdx.x], lds[k][threadIdx.x + s]);
            for (int k = 3; k < 6; k++) lds[k][threadIdx.x] = max(lds[k][threadIdx.x], lds[k][threadIdx.x + s]);
        }
        __syncthreads();
    }
    if (threadIdx.x < 6) bbox[threadIdx.x] = lds[threadIdx.x][0];
}

__global__ void k_init_bbox(int *bbox) {
    if (threadIdx.x < 3) bbox[threadIdx.x] = 2147483647;
    else if (threadIdx.x < 6) bbox[threadIdx.x] = -2147483647 - 1;
}

// ------------------------------------------------------------------------
// spatial-hash collision guard: two cells within Chebyshev distance 2c that
// share a bucket key would make the reference visit a bucket twice
// (spatial_hashing.rs:61-103); flags such a grid so the host can refuse it.
// ------------------------------------------------------------------------
__device__ __forceinline__ unsigned long long map_u(int k) {
    return k >= 0 ? (unsigned long long)k * 2ull : (unsigned long long)(-(long long)k) * 2ull + 1ull;
}
__device__ __forceinline__ unsigned long long hash3(int k, int l, int m) {
    return (map_u(k) * 73856093ull) ^ (map_u(l) * 19349663ull) ^ (map_u(m) * 83492791ull);
}

__global__ void k_hash_check(GridDev g, int reach, unsigned int *flag) {
    size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t nc = (size_t)g.dim[0] * g.dim[1] * g.dim[2];
    if (t >= nc) return;
    int gz = (int)(t % g.dim[2]);
    int gy = (int)((t / g.dim[2]) % g.dim[1]);
    int gx = (int)(t / ((size_t)g.dim[2] * g.dim[1]));
    int cx = gx + g.org[0], cy = gy + g.org[1], cz = gz + g.org[2];
    unsigned long long h0 = hash3(cx, cy, cz);
    for (int a = 0; a <= reach; a++)
        for (int b = -reach; b <= reach; b++)
            for (int c = -reach; c <= reach; c++) {
                if (a == 0 && (b < 0 || (b == 0 && c <= 0))) continue;
                if (hash3(cx + a, cy + b, cz + c) == h0) atomicOr(flag, 1u);
            }
}

// The bucket partners (GridDev::part) of every cell of g's box widened by
// g.pm: the cells of the box at most `span` away per axis whose keys equal
// the cell's own (kNoCell where there are fewer than two). flags[0] is set
// when any cell has a partner, flags[1] when one has more than two.
__global__ void k_hash_partners(GridDev g, int span, uint32_t *part, unsigned int *flags) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t e0 = (size_t)g.dim[0] + 2 * g.pm, e1 = (size_t)g.dim[1] + 2 * g.pm,
                 e2 = (size_t)g.dim[2] + 2 * g.pm;
    if (t >= e0 * e1 * e2) return;
    // the cell's coordinates relative to the box origin
    const int gz = (int)(t % e2) - g.pm;
    const int gy = (int)((t / e2) % e1) - g.pm;
    const int gx = (int)(t / (e2 * e1)) - g.pm;
    const unsigned long long h0 = hash3(gx + g.org[0], gy + g.org[1], gz + g.org[2]);
    uint32_t found[2] = {kNoCell, kNoCell};
    int nf = 0;
    for (int a = -span; a <= span; a++) {
        const int bx = gx + a;
        if (bx < 0 || bx >= g.dim[0]) continue;
        for (int b = -span; b <= span; b++) {
            const int by = gy + b;
            if (by < 0 || by >= g.dim[1]) continue;
            for (int c = -span; c <= span; c++) {
                const int bz = gz + c;
                if (bz < 0 || bz >= g.dim[2] || (a == 0 && b == 0 && c == 0)) continue;
                if (hash3(bx + g.org[0], by + g.org[1], bz + g.org[2]) != h0) continue;
                if (nf < 2)
                    found[nf] = ((uint32_t)bx * (uint32_t)g.dim[1] + (uint32_t)by) * (uint32_t)g.dim[2] +
                                (uint32_t)bz;
                nf++;
            }
        }
    }
    part[2 * t] = found[0];
    part[2 * t + 1] = found[1];
    if (nf > 0) atomicOr(&flags[0], 1u);
    if (nf > 2) atomicOr(&flags[1], 1u);
}

// ------------------------------------------------------------------------
// counting sort by (cell key, reference index)
// ------------------------------------------------------------------------
__device__ __forceinline__ uint32_t cell_key(const GridDev &g, int cx, int cy, int cz) {
    return ((uint32_t)(cx - g.org[0]) * (uint32_t)g.dim[1] + (uint32_t)(cy - g.org[1])) *
               (uint32_t)g.dim[2] +
           (uint32_t)(cz - g.org[2]);
}

__global__ void k_cell_count(size_t n, const int *ci, GridDev g, uint32_t *count) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    atomicAdd(&count[cell_key(g, ci[3 * i], ci[3 * i + 1], ci[3 * i + 2])], 1u);
}

__global__ void k_cell_scatter(size_t n, const int *ci, GridDev g, const uint32_t *start,
                               uint32_t *cursor, uint32_t *perm) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t k = cell_key(g, ci[3 * i], ci[3 * i + 1], ci[3 * i + 2]);
    uint32_t slot = start[k] + atomicAdd(&cursor[k], 1u);
    perm[slot] = (uint32_t)i;
}

// order each cell's members by reference index (the reference's bucket order)
__global__ void k_cell_order(size_t ncell, const uint32_t *start, uint32_t *perm,
                             const uint32_t *ref) {
    size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncell) return;
    uint32_t lo = start[c], hi = start[c + 1];
    for (uint32_t a = lo + 1; a < hi; a++) {
        uint32_t v = perm[a];
        uint32_t key = ref ? ref[v] : v;
        uint32_t b = a;
        while (b > lo) {
            uint32_t u = perm[b - 1];
            if ((ref ? ref[u] : u) <= key) break;
            perm[b] = u;
            b--;
        }
        perm[b] = v;
    }
}

template <class T>
__global__ void k_gather(size_t n, const uint32_t *perm, const T *src, T *dst) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[perm[i]];
}
template __global__ void k_gather<double>(size_t, const uint32_t *, const double *, double *);

template <class T>
__global__ void k_scatter(size_t n, const uint32_t *perm, const T *src, T *dst) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[perm[i]] = src[i];
}
template __global__ void k_scatter<double>(size_t, const uint32_t *, const double *, double *);
template __global__ void k_gather<uint32_t>(size_t, const uint32_t *, const uint32_t *,
                                             uint32_t *);
template __global__ void k_gather<int>(size_t, const uint32_t *, const int *, int *);

// sorted structure-of-arrays copy of [n][3] positions: dst[i] = src[perm[i]]
__global__ void k_gather_aos(size_t n, const uint32_t *perm, const double *p, double *x,
                             double *y, double *z) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const size_t j = perm[i];
    x[i] = p[3 * j];
    y[i] = p[3 * j + 1];
    z[i] = p[3 * j + 2];
}

// ------------------------------------------------------------------------
// neighbor search (spatial_hashing.rs:75-103): for every query, the samples
// of the (2c+1)^3 stencil with |x_q - x_s| < range. The stencil is walked in
// the reference's (dx, dy, dz) order; with z the fastest key axis each
// (dx, dy) row is one contiguous range of the sorted samples, and members of
// a cell are ordered by reference index, so the emitted order equals the
// reference's. The predicate d2 < dthr, with dthr the least double whose
// correctly rounded sqrt is >= range, is exactly sqrt(d2) < range.
// ------------------------------------------------------------------------
// One query's stencil walk; emit(cnt, j) receives the cnt-th accepted sample
// (sorted index j). Returns the number accepted.
template <class Emit>
__device__ __forceinline__ uint32_t stencil_walk(double px, double py, double pz, double cs,
                                                 int reach, double dthr, const GridDev &g,
                                                 const double *sx, const double *sy,
                                                 const double *sz, Emit emit) {
    const int cx = sat_i32(floor(px / cs)), cy = sat_i32(floor(py / cs)),
              cz = sat_i32(floor(pz / cs));
    uint32_t cnt = 0;
    if (g.part) {  // shared hash buckets: the reference's bucket visits
        bucket_walk(g, cx, cy, cz, reach, [&](uint32_t j) {
            double ddx = px - sx[j], ddy = py - sy[j], ddz = pz - sz[j];
            double d2 = (ddx * ddx + ddy * ddy) + ddz * ddz;
            if (d2 < dthr) {
                emit(cnt, j);
                cnt++;
            }
        });
        return cnt;
    }
    int zlo = cz - reach - g.org[2], zhi = cz + reach - g.org[2];
    if (zlo < 0) zlo = 0;
    if (zhi > g.dim[2] - 1) zhi = g.dim[2] - 1;
    if (zlo > zhi) return 0;
    for (int dx = -reach; dx <= reach; dx++) {
        int gx = cx + dx - g.org[0];
        if (gx < 0 || gx >= g.dim[0]) continue;
        for (int dy = -reach; dy <= reach; dy++) {
            int gy = cy + dy - g.org[1];
            if (gy < 0 || gy >= g.dim[1]) continue;
            uint32_t row = ((uint32_t)gx * (uint32_t)g.dim[1] + (uint32_t)gy) * (uint32_t)g.dim[2];
            uint32_t a = g.start[row + zlo], b = g.start[row + zhi + 1];
            for (uint32_t j = a; j < b; j++) {
                double ddx = px - sx[j], ddy = py - sy[j], ddz = pz - sz[j];
                double d2 = (ddx * ddx + ddy * ddy) + ddz * ddz;
                if (d2 < dthr) {
                    emit(cnt, j);
                    cnt++;
                }
            }
        }
    }
    return cnt;
}

template <bool FILL>
__global__ void k_neighbors(size_t nq, const double *qx, const double *qy, const double *qz,
                            double cs, int reach, double dthr, GridDev g, const double *sx,
                            const double *sy, const double *sz, const uint32_t *off,
                            uint32_t *out_count, uint32_t *out_idx) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nq) return;
    const uint32_t o = FILL ? off[i] : 0;
    const uint32_t cnt = stencil_walk(qx[i], qy[i], qz[i], cs, reach, dthr, g, sx, sy, sz,
                                      [&](uint32_t k, uint32_t j) {
                                          if (FILL) out_idx[o + k] = j;
                                      });
    if (!FILL) out_count[i] = cnt;
}
template __global__ void k_neighbors<false>(size_t, const double *, const double *,
                                            const double *, double, int, double, GridDev,
                                            const double *, const double *, const double *,
                                            const uint32_t *, uint32_t *, uint32_t *);
template __global__ void k_neighbors<true>(size_t, const double *, const double *,
                                           const double *, double, int, double, GridDev,
                                           const double *, const double *, const double *,
                                           const uint32_t *, uint32_t *, uint32_t *);

// [nq][3] queries; FILL writes the caller's sample index perm[j] as uint64_t
template <bool FILL>
__global__ void k_neighbors_aos(size_t nq, const double *q, double cs, int reach, double dthr,
                                GridDev g, const double *sx, const double *sy, const double *sz,
                                const uint32_t *off, uint32_t *out_count, const uint32_t *perm,
                                uint64_t *out_idx) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nq) return;
    const uint32_t o = FILL ? off[i] : 0;
    const uint32_t cnt = stencil_walk(q[3 * i], q[3 * i + 1], q[3 * i + 2], cs, reach, dthr, g,
                                      sx, sy, sz, [&](uint32_t k, uint32_t j) {
                                          if (FILL) out_idx[o + k] = (uint64_t)perm[j];
                                      });
    if (!FILL) out_count[i] = cnt;
}
template __global__ void k_neighbors_aos<false>(size_t, const double *, double, int, double,
                                                GridDev, const double *, const double *,
                                                const double *, const uint32_t *, uint32_t *,
                                                const uint32_t *, uint64_t *);
template __global__ void k_neighbors_aos<true>(size_t, const double *, double, int, double,
                                               GridDev, const double *, const double *,
                                               const double *, const uint32_t *, uint32_t *,
                                               const uint32_t *, uint64_t *);

// ------------------------------------------------------------------------
// exclusive scan of uint32 (three-phase, 1024 elements per block)
// ------------------------------------------------------------------------
__global__ void k_scan_block(const uint32_t *in, size_t n, uint32_t *out, uint32_t *block_sums) {
    __shared__ uint32_t s[1024];
    const int t = threadIdx.x;  // 512 threads, 2 elements each
    size_t base = (size_t)blockIdx.x * 1024;
    uint32_t a = base + 2 * t < n ? in[base + 2 * t] : 0u;
    uint32_t b = base + 2 * t + 1 < n ? in[base + 2 * t + 1] : 0u;
    s[2 * t] = a;
    s[2 * t + 1] = b;
    __syncthreads();
    for (int d = 1; d < 1024; d <<= 1) {  // Hillis-Steele on pairs
        int idx = (t + 1) * 2 * d - 1;
        if (idx < 1024) s[idx] += s[idx - d];
        __syncthreads();
    }
    if (t == 0) {
        if (block_sums) block_sums[blockIdx.x] = s[1023];
        s[1023] = 0;
    }
    __syncthreads();
    for (int d = 512; d >= 1; d >>= 1) {
        int idx = (t + 1) * 2 * d - 1;
        if (idx < 1024) {
            uint32_t tmp = s[idx - d];
            s[idx - d] = s[idx];
            s[idx] += tmp;
        }
        __syncthreads();
    }
    if (base + 2 * t < n) out[base + 2 * t] = s[2 * t];
    if (base + 2 * t + 1 < n) out[base + 2 * t + 1] = s[2 * t + 1];
}

__global__ void k_scan_add(uint32_t *out, size_t n, const uint32_t *block_offsets) {
    size_t i = (size_t)blockIdx.x * 1024 + threadIdx.x;
    uint32_t add = block_offsets[blockIdx.x];
    if (i < n) out[i] += add;
    if (i + 512 < n && threadIdx.x + 512 < 1024) out[i + 512] += add;
}

// StaticSample pseudo volume V_b = gamma_V / sum_k W(x_k - x_b)
// (static_sample_boundary.rs:354-390), samples in sorted order
__global__ void k_static_volume(size_t n, KernelConst K, double gamma_v, const double *x,
                                const double *y, const double *z, const uint32_t *off,
                                const uint32_t *idx, double *vol) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double inv = 0.;
    for (uint32_t e = off[i]; e < off[i + 1]; e++) {
        uint32_t j = idx[e];
        inv += kernel_fn(K, x[j] - x[i], y[j] - y[i], z[j] - z[i]);
    }
    vol[i] = gamma_v / inv;
}

// neighbor indices in sorted sample space -> caller's sample index (u64)
__global__ void k_map_indices(size_t n, const uint32_t *perm, const uint32_t *in, uint64_t *out) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = perm ? (uint64_t)perm[in[i]] : (uint64_t)in[i];
}

// out[n] = out[n-1] + in[n-1]: the total after an exclusive scan of n > 0 counts
__global__ void k_scan_total(const uint32_t *in, size_t n, uint32_t *out) {
    if (blockIdx.x == 0 && threadIdx.x == 0) out[n] = out[n - 1] + in[n - 1];
}

__global__ void k_widen_u32(size_t n, const uint32_t *in, uint64_t *out) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = (uint64_t)in[i];
}

}  // namespace sph

namespace sph {

// ------------------------------------------------------------------------
// Volume-map construction (DiscretizedBoundaryFields::new,
// volume_map_boundary.rs:1593-1624): the node values of the volume map on
// the lattice (VmGrid) from a discretized signed-distance field
// ------------------------------------------------------------------------
__device__ __forceinline__ bool vm_is_node(int ix, int iy, int iz) {
    // serendipity nodes: corners (all offsets 0 mod 3) and edge nodes (one
    // offset 1 or 2 mod 3)
    return ((ix % 3 != 0) + (iy % 3 != 0) + (iz % 3 != 0)) <= 1;
}

// cubic_extension_fn (volume_map_boundary.rs:1716-1726)
__device__ __forceinline__ double vm_cubic_extension(double sd, double h, double wfac) {
    if (sd >= h) return 0.;
    if (sd <= 0.) return 1.;
    const double q = sqrt(dot3(sd, 0., 0., sd, 0., 0.)) / h;
    const double q0 = sqrt(dot3(0., 0., 0., 0., 0., 0.)) / h;
    return (wfac * kernel_w(q)) / (wfac * kernel_w(q0));
}

__global__ void k_vm_classify(LatticeDev sdf, VmGrid g, double h, double bv, double *values,
                              uint8_t *status, uint32_t *band, unsigned int *nband) {
    const size_t nx = 3 * (size_t)g.n[0] + 1, ny = 3 * (size_t)g.n[1] + 1,
                 nz = 3 * (size_t)g.n[2] + 1;
    const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nx * ny * nz) return;
    const int ix = (int)(k % nx), iy = (int)((k / nx) % ny), iz = (int)(k / (nx * ny));
    values[k] = 0.0;
    if (!vm_is_node(ix, iy, iz)) {
        status[k] = 0;
        return;
    }
    // x_min + key / 3 * dx (discretization.rs:171-175)
    const double px = g.x_min[0] + (double)ix / 3.0 * g.dx,
                 py = g.x_min[1] + (double)iy / 3.0 * g.dx,
                 pz = g.x_min[2] + (double)iz / 3.0 * g.dx;
    double sd;
    status[k] = 1;
    if (serendipity(sdf, px, py, pz, sd)) return;  // outside the band: 0 (SDFnWrapper::volume)
    if (sd >= 2.0 * h) return;
    if (sd <= -h) {
        values[k] = 4.0 / 3.0 * 3.14159265358979323846 * ((h * h) * h);
        return;
    }
    status[k] = 2;  // quadrature pending
    band[atomicAdd(nband, 1u)] = (uint32_t)k;
}

// gauss_legendre_integrate_ball (utilities/discretization.rs:33-79) of the
// cubic extension of the SDF around one band node per workgroup: thread t
// sums the phi rule for (r_i, theta_j) pairs, then the theta and r rules are
// summed in node order, as nested sequential sums
__global__ void k_vm_quadrature(LatticeDev sdf, VmGrid g, VmQuad Q, const uint32_t *band,
                                double *values, uint8_t *status) {
    __shared__ double tj[kVmMaxDegree][kVmMaxDegree + 1];
    __shared__ double ti[kVmMaxDegree];
    __shared__ int err;
    const double PI = 3.14159265358979323846;
    const size_t nx = 3 * (size_t)g.n[0] + 1, ny = 3 * (size_t)g.n[1] + 1;
    const uint32_t k = band[blockIdx.x];
    const int ix = (int)(k % nx), iy = (int)((k / nx) % ny), iz = (int)(k / (nx * ny));
    const double px = g.x_min[0] + (double)ix / 3.0 * g.dx,
                 py = g.x_min[1] + (double)iy / 3.0 * g.dx,
                 pz = g.x_min[2] + (double)iz / 3.0 * g.dx;
    const int D = Q.degree;
    if (threadIdx.x == 0) err = 0;
    __syncthreads();
    for (int ij = threadIdx.x; ij < D * D; ij += blockDim.x) {
        const int i = ij / D, j = ij % D;
        const double r = Q.r[i], st = Q.st[j], ct = Q.ct[j];
        double acc = 0.0;
        int e = 0;
        for (int kk = 0; kk < D; kk++) {
            double sd;
            e = serendipity(sdf, px + r * st * Q.cp[kk], py + r * st * Q.sp[kk], pz + r * ct, sd);
            if (e) break;
            acc += vm_cubic_extension(sd, Q.h, Q.wfac) * Q.w[kk];
        }
        if (e) err = 1;
        tj[i][j] = st * (0.5 * (2. * PI - 0.0) * acc) * Q.w[j];
    }
    __syncthreads();
    if ((int)threadIdx.x < D) {
        const int i = threadIdx.x;
        double acc = 0.0;
        for (int j = 0; j < D; j++) acc += tj[i][j];
        ti[i] = (Q.r[i] * Q.r[i]) * (0.5 * (PI - 0.0) * acc) * Q.w[i];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        if (err) {
            status[k] = 3;  // an SDF evaluation failed: the node has no value
            return;
        }
        double acc = 0.0;
        for (int i = 0; i < D; i++) acc += ti[i];
        double v = 0.5 * (Q.h - 0.0) * acc;
        v = v < 0.0 ? 0.0 : (v > Q.bv ? Q.bv : v);  // clamp(0, ball_volume)
        values[k] = v;
        status[k] = 1;
    }
}

// cells whose 32 nodes all have values are kept (prune lower bound 0 removes
// nothing else, values being >= 0); their nodes are present
__global__ void k_vm_cells(VmGrid g, const uint8_t *status, uint8_t *present) {
    const size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t ncell = (size_t)g.n[0] * g.n[1] * g.n[2];
    if (c >= ncell) return;
    const size_t nx = 3 * (size_t)g.n[0] + 1, ny = 3 * (size_t)g.n[1] + 1;
    const size_t cx = c % g.n[0], cy = (c / g.n[0]) % g.n[1], cz = c / ((size_t)g.n[0] * g.n[1]);
    const size_t base = (3 * cz * ny + 3 * cy) * nx + 3 * cx;
    for (int i = 0; i < 32; i++)
        if (status[base + ((size_t)kNodeOff[i][2] * ny + kNodeOff[i][1]) * nx + kNodeOff[i][0]] != 1)
            return;
    for (int i = 0; i < 32; i++)
        present[base + ((size_t)kNodeOff[i][2] * ny + kNodeOff[i][1]) * nx + kNodeOff[i][0]] = 1;
}

__global__ void k_vm_finish(size_t n, const uint8_t *present, double *values) {
    const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n && !present[k]) values[k] = 0.0;
}

}  // namespace sph

namespace sph {

// ------------------------------------------------------------------------
// Quantities at arbitrary positions (SPHSystem::get_quantity_at_positions,
// sph/mod.rs:241-341; quantities.rs:14-330). One thread per query walks the
// fluid's (2c+1)^3 stencil in the order of k_neighbors, so every sum takes
// its terms in the reference's neighbor order; boundaries contribute to the
// speed and the volume only, as in the reference.
// ------------------------------------------------------------------------
template <class Fn>
__device__ __forceinline__ void stencil_walk(const GridDev &g, int cx, int cy, int cz, int reach,
                                             Fn &&fn) {
    if (g.part) {  // shared hash buckets: the reference's bucket visits
        bucket_walk(g, cx, cy, cz, reach, fn);
        return;
    }
    int zlo = cz - reach - g.org[2], zhi = cz + reach - g.org[2];
    if (zlo < 0) zlo = 0;
    if (zhi > g.dim[2] - 1) zhi = g.dim[2] - 1;
    if (zlo > zhi) return;
    for (int dx = -reach; dx <= reach; dx++) {
        const int gx = cx + dx - g.org[0];
        if (gx < 0 || gx >= g.dim[0]) continue;
        for (int dy = -reach; dy <= reach; dy++) {
            const int gy = cy + dy - g.org[1];
            if (gy < 0 || gy >= g.dim[1]) continue;
            const uint32_t row = ((uint32_t)gx * (uint32_t)g.dim[1] + (uint32_t)gy) * (uint32_t)g.dim[2];
            const uint32_t a = g.start[row + zlo], b = g.start[row + zhi + 1];
            for (uint32_t j = a; j < b; j++) fn(j);
        }
    }
}

template <int Q>
__global__ void k_quantity(size_t nq, const double *qx, const double *qy, const double *qz,
                           QueryConst C, QueryFluid F, int nb, const BodyDev *bodies,
                           const GridDev *sgrid, float *out) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nq) return;
    const double px = qx[i], py = qy[i], pz = qz[i];
    const int cx = sat_i32(floor(px / C.cs)), cy = sat_i32(floor(py / C.cs)),
              cz = sat_i32(floor(pz / C.cs));
    const KernelConst K = C.K;
    double acc = 0.0, ax = 0.0, ay = 0.0, az = 0.0;
    stencil_walk(F.g, cx, cy, cz, C.reach, [&](uint32_t j) {
        const uint32_t s = F.perm[j];
        const double rx = px - F.x[s], ry = py - F.y[s], rz = pz - F.z[s];
        const double d2 = (rx * rx + ry * ry) + rz * rz;
        if (!(d2 < C.dthr)) return;
        const double w = K.wfac * kernel_w(sqrt(d2) / K.h);
        if (Q == kQSpeed) {  // get_speed (quantities.rs:64-120)
            const double Vs = F.V[s];
            ax += (F.vx[s] * Vs) * w;
            ay += (F.vy[s] * Vs) * w;
            az += (F.vz[s] * Vs) * w;
        } else if (Q == kQVolume) {  // get_volume (:14-61)
            acc += C.V0 * w;
        } else if (Q == kQDensity) {  // get_density (:123-170)
            acc += F.m[s] * w;
        } else if (Q == kQDensityError) {
            // get_density_error (:173-228) over the masses its caller passes
            // as sample volumes (sph/mod.rs:297-306)
            const double ms = F.m[s];
            if (ms < C.V0) acc += ((C.V0 / ms - 1.) * ms) * w;
        } else if (Q == kQPressure) {  // get_pressure (:231-280)
            acc += (F.p[s] * F.V[s]) * w;
        } else {  // get_kinetic_energy (:283-330)
            const double ux = F.vx[s], uy = F.vy[s], uz = F.vz[s];
            acc += (((0.5 * F.m[s]) * ((ux * ux + uy * uy) + uz * uz)) * F.V[s]) * w;
        }
    });
    if (Q == kQSpeed || Q == kQVolume) {
        for (int bi = 0; bi < nb; bi++) {
            const BodyDev &b = bodies[bi];
            if (b.kind == 0) {  // static samples (velocity zero unless on a rigid body)
                stencil_walk(sgrid[bi], cx, cy, cz, C.reach, [&](uint32_t j) {
                    const double rx = px - b.sx[j], ry = py - b.sy[j], rz = pz - b.sz[j];
                    const double d2 = (rx * rx + ry * ry) + rz * rz;
                    if (!(d2 < C.dthr)) return;
                    const double w = K.wfac * kernel_w(sqrt(d2) / K.h);
                    if (Q == kQVolume) {
                        acc += b.svol[j] * w;
                    } else {  // samples move only with a rigid body
                        ax += ((b.dynamic ? b.svx[j] : 0.0) * b.svol[j]) * w;
                        ay += ((b.dynamic ? b.svy[j] : 0.0) * b.svol[j]) * w;
                        az += ((b.dynamic ? b.svz[j] : 0.0) * b.svol[j]) * w;
                    }
                });
            } else {
                double ox, oy, oz, vol;
                if (!vm_normal_sample(b, px, py, pz, K.h, C.dxr, C.vmax, ox, oy, oz, vol)) continue;
                const double w = kernel_fn(K, px - ox, py - oy, pz - oz);
                if (Q == kQVolume) {
                    acc += vol * w;
                } else {
                    double vx, vy, vz;
                    body_velocity(b, ox, oy, oz, vx, vy, vz);
                    ax += (vx * vol) * w;
                    ay += (vy * vol) * w;
                    az += (vz * vol) * w;
                }
            }
        }
    }
    double res;
    if (Q == kQSpeed)
        res = sqrt((ax * ax + ay * ay) + az * az);
    else if (Q == kQVolume)
        res = C.V0 / acc;
    else if (Q == kQDensityError)
        res = 100. * acc;
    else
        res = acc;
    out[i] = (float)res;
}

#define INSTANTIATE_QUANTITY(Q)                                                                  \
    template __global__ void k_quantity<Q>(size_t, const double *, const double *,              \
                                           const double *, QueryConst, QueryFluid, int,          \
                                           const BodyDev *, const GridDev *, float *);
INSTANTIATE_QUANTITY(kQSpeed)
INSTANTIATE_QUANTITY(kQVolume)
INSTANTIATE_QUANTITY(kQDensity)
INSTANTIATE_QUANTITY(kQDensityError)
INSTANTIATE_QUANTITY(kQPressure)
INSTANTIATE_QUANTITY(kQKineticEnergy)
#undef INSTANTIATE_QUANTITY

// ------------------------------------------------------------------------
// Rigid-rigid contact cells: one workgroup per candidate sub-cell
// ------------------------------------------------------------------------
// cubic_extension_fn_derivative (volume_map_boundary.rs:1734-1741): the x
// component of grad W at (sd, 0, 0) over W(0)
__device__ __forceinline__ double contact_ext_derivative(const KernelConst &K, double sd) {
    if (sd >= K.h || sd <= 0.0) return 0.0;
    double gx, gy, gz;
    kernel_grad(K, sd, 0.0, 0.0, gx, gy, gz);
    return gx / kernel_fn(K, 0.0, 0.0, 0.0);
}

// The candidate test of generate_contact_cells (:124-146) at the sub-cell
// centre, then build_contact_cell (:702-768) and the sums of
// compute_diagonal (:835-882) over the degree^3 Gauss-Legendre points
// (quadrature_points, :805-830). Each thread evaluates one point per round
// and leaves its terms in LDS (+0 for a point that is skipped, which leaves
// a sum that starts at +0 unchanged); threads 0..16 then add one term
// column each in point order, so every sum is the reference's sequential
// sum over the kept points. Active cells take a slot from `nact` and are
// stored while the slot is below `cap`; the host orders them by candidate
// index.
__global__ __launch_bounds__(kContactBlock) void k_contact_cells(
    const ContactPairDev *pairs, int npairs, const BodyDev *bodies, ContactQuadDev Q,
    uint32_t base, unsigned int *nact, uint32_t cap, uint32_t *cell_idx, double *sums) {
    __shared__ double terms[kContactBlock][kContactSums];
    __shared__ double total[kContactSums];
    const uint32_t c = base + blockIdx.x;
    int pi = 0;
    while (pi + 1 < npairs && pairs[pi + 1].first <= c) pi++;
    const ContactPairDev &P = pairs[pi];
    const BodyDev &A = bodies[P.a];
    const BodyDev &B = bodies[P.b];
    const uint32_t local = c - P.first;
    const uint32_t ix = local % P.nn[0], iy = (local / P.nn[0]) % P.nn[1],
                   iz = local / (P.nn[0] * P.nn[1]);
    const double fx = P.omn[0] + (double)ix * P.cs[0], fy = P.omn[1] + (double)iy * P.cs[1],
                 fz = P.omn[2] + (double)iz * P.cs[2];
    const double tx = fx + P.cs[0], ty = fy + P.cs[1], tz = fz + P.cs[2];
    const double hx = (tx - fx) * 0.5, hy = (ty - fy) * 0.5, hz = (tz - fz) * 0.5;
    const double cx = fx + hx, cy = fy + hy, cz = fz + hz;
    {
        const double dx = tx - fx, dy = ty - fy, dz = tz - fz;
        const double bound = 2.0 * (sqrt(dot3(dx, dy, dz, dx, dy, dz)) * 0.5);
        double pa, pb;
        // uniform over the workgroup: every thread evaluates the same centre
        if (body_sd(A, cx, cy, cz, pa) || body_sd(B, cx, cy, cz, pb)) return;
        if (!(pa - bound < 0. && pb - bound < 0.)) return;
    }
    const KernelConst &K = Q.K;
    const int n = Q.degree, npts = n * n * n;
    const int t = threadIdx.x;
    double acc = 0.0;
    for (int r0 = 0; r0 < npts; r0 += kContactBlock) {
        double v[kContactSums];
#pragma unroll
        for (int k = 0; k < kContactSums; k++) v[k] = 0.0;
        const int q = r0 + t;
        if (q < npts) {
            const int i = q / (n * n), j = (q / n) % n, k = q % n;
            const double px = cx + Q.x[i] * hx, py = cy + Q.x[j] * hy, pz = cz + Q.x[k] * hz;
            const double w = Q.w[i] * Q.w[j] * Q.w[k] * hx * hy * hz;
            double pa, pb, gax, gay, gaz, gbx, gby, gbz;
            if (!body_sd(A, px, py, pz, pa) && !body_sd(B, px, py, pz, pb)) {
                const double ua = 2.0 * pa + K.h, ub = 2.0 * pb + K.h;
                const double sa = vm_cubic_extension(ua, K.h, K.wfac);
                const double sb = vm_cubic_extension(ub, K.h, K.wfac);
                if (sa > 0.0 && sb > 0.0 && !body_sdg(A, px, py, pz, gax, gay, gaz) &&
                    !body_sdg(B, px, py, pz, gbx, gby, gbz)) {
                    const double spa = 2.0 * contact_ext_derivative(K, ua);
                    const double spb = 2.0 * contact_ext_derivative(K, ub);
                    const double wg = w * sa * sb;
                    v[0] = wg;
                    v[1] = wg * px;
                    v[2] = wg * py;
                    v[3] = wg * pz;
                    const double wa = w * spa * sb;
                    v[4] = wa * gax;
                    v[5] = wa * gay;
                    v[6] = wa * gaz;
                    if (A.dynamic) {
                        double ox, oy, oz;
                        cross3(px - A.com[0], py - A.com[1], pz - A.com[2], gax, gay, gaz, ox,
                               oy, oz);
                        v[7] = wa * ox;
                        v[8] = wa * oy;
                        v[9] = wa * oz;
                    }
                    const double wb = w * spb * sa;
                    v[10] = wb * gbx;
                    v[11] = wb * gby;
                    v[12] = wb * gbz;
                    if (B.dynamic) {
                        double ox, oy, oz;
                        cross3(px - B.com[0], py - B.com[1], pz - B.com[2], gbx, gby, gbz, ox,
                               oy, oz);
                        v[13] = wb * ox;
                        v[14] = wb * oy;
                        v[15] = wb * oz;
                    }
                    v[16] = 1.0;
                }
            }
        }
#pragma unroll
        for (int k = 0; k < kContactSums; k++) terms[t][k] = v[k];
        __syncthreads();
        if (t < kContactSums) {
            const int lim = min(kContactBlock, npts - r0);
            for (int u = 0; u < lim; u++) acc += terms[u][t];
        }
        __syncthreads();
    }
    if (t < kContactSums) total[t] = acc;
    __syncthreads();
    if (t == 0 && total[0] > 0.0 && total[16] > 0.0) {  // v_c > 0 with points kept
        const unsigned int slot = atomicAdd(nact, 1u);
        if (slot < cap) {  // the host reruns with room for every active cell
            cell_idx[slot] = c;
            for (int k = 0; k < kContactSums; k++)
                sums[(size_t)slot * kContactSums + k] = total[k];
        }
    }
}

// ------------------------------------------------------------------------
// Signed distance to an oriented triangle mesh at the lattice nodes
// ------------------------------------------------------------------------
__device__ __forceinline__ void sub3v(const double *a, const double *b, double *o) {
    o[0] = a[0] - b[0];
    o[1] = a[1] - b[1];
    o[2] = a[2] - b[2];
}
__device__ __forceinline__ double dot3v(const double *a, const double *b) {
    return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2];
}

// The closest point q of triangle (a, b, c) to p by Voronoi-region tests
// (Ericson, Real-Time Collision Detection, 5.1.5) and its feature: 0..2 the
// vertices a, b, c; 3 edge ab, 4 edge bc, 5 edge ca; 6 the face
__device__ __forceinline__ int mesh_closest(const double *p, const double *a, const double *b,
                                            const double *c, double *q) {
    double ab[3], ac[3], ap[3], bp[3], cp[3];
    sub3v(b, a, ab);
    sub3v(c, a, ac);
    sub3v(p, a, ap);
    const double d1 = dot3v(ab, ap), d2 = dot3v(ac, ap);
    if (d1 <= 0.0 && d2 <= 0.0) {
        q[0] = a[0], q[1] = a[1], q[2] = a[2];
        return 0;
    }
    sub3v(p, b, bp);
    const double d3 = dot3v(ab, bp), d4 = dot3v(ac, bp);
    if (d3 >= 0.0 && d4 <= d3) {
        q[0] = b[0], q[1] = b[1], q[2] = b[2];
        return 1;
    }
    const double vc = d1 * d4 - d3 * d2;
    if (vc <= 0.0 && d1 >= 0.0 && d3 <= 0.0) {
        const double s = d1 / (d1 - d3);
        for (int k = 0; k < 3; k++) q[k] = a[k] + s * ab[k];
        return 3;
    }
    sub3v(p, c, cp);
    const double d5 = dot3v(ab, cp), d6 = dot3v(ac, cp);
    if (d6 >= 0.0 && d5 <= d6) {
        q[0] = c[0], q[1] = c[1], q[2] = c[2];
        return 2;
    }
    const double vb = d5 * d2 - d1 * d6;
    if (vb <= 0.0 && d2 >= 0.0 && d6 <= 0.0) {
        const double s = d2 / (d2 - d6);
        for (int k = 0; k < 3; k++) q[k] = a[k] + s * ac[k];
        return 5;
    }
    const double va = d3 * d6 - d5 * d4;
    if (va <= 0.0 && (d4 - d3) >= 0.0 && (d5 - d6) >= 0.0) {
        const double s = (d4 - d3) / ((d4 - d3) + (d5 - d6));
        for (int k = 0; k < 3; k++) q[k] = b[k] + s * (c[k] - b[k]);
        return 4;
    }
    const double den = 1.0 / ((va + vb) + vc);
    const double s = vb * den, w = vc * den;
    for (int k = 0; k < 3; k++) q[k] = (a[k] + ab[k] * s) + ac[k] * w;
    return 6;
}

// One thread per lattice position; serendipity nodes take the distance to
// the closest point over all triangles, signed by the pseudo-normal of its
// feature. The least (distance, mesh index) wins, which is the first closest
// triangle in mesh order whatever order the triangles are visited in.
// Triangles come in spatially compact tiles staged through LDS; a node skips
// a tile whose bounding box lies farther than its best distance so far (with
// a margin over the rounding of both distances, so that no skipped triangle
// could have won), and the workgroup skips the tile when all its nodes do.
__global__ __launch_bounds__(256) void k_mesh_sdf(MeshDev M, VmGrid g, double *values) {
    __shared__ double tile[kMeshTile * 9];
    __shared__ uint32_t tile_orig[kMeshTile];
    const size_t nx = 3 * (size_t)g.n[0] + 1, ny = 3 * (size_t)g.n[1] + 1,
                 nz = 3 * (size_t)g.n[2] + 1;
    const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool in = k < nx * ny * nz;
    const int ix = in ? (int)(k % nx) : 0, iy = in ? (int)((k / nx) % ny) : 0,
              iz = in ? (int)(k / (nx * ny)) : 0;
    const bool node = in && vm_is_node(ix, iy, iz);
    const double p[3] = {g.x_min[0] + (double)ix / 3.0 * g.dx, g.x_min[1] + (double)iy / 3.0 * g.dx,
                         g.x_min[2] + (double)iz / 3.0 * g.dx};
    double best = __builtin_inf(), bq[3] = {0.0, 0.0, 0.0};
    uint32_t bt = 0xffffffffu;
    int bf = -1;
    for (uint32_t base = 0; base < M.nt; base += kMeshTile) {
        const uint32_t cnt = min((uint32_t)kMeshTile, M.nt - base);
        const double *bx = M.box + 6 * (size_t)(base / kMeshTile);
        double bd2 = 0.0;
        for (int d = 0; d < 3; d++) {
            const double gap = fmax(fmax(bx[d] - p[d], p[d] - bx[3 + d]), 0.0);
            bd2 += gap * gap;
        }
        const bool need = node && !(bd2 > best + best * 1e-6 + M.slack);
        if (!__syncthreads_or(need)) continue;
        for (uint32_t i = threadIdx.x; i < 9 * cnt; i += blockDim.x)
            tile[i] = M.tri[(size_t)base * 9 + i];
        for (uint32_t i = threadIdx.x; i < cnt; i += blockDim.x) tile_orig[i] = M.orig[base + i];
        __syncthreads();
        if (!need) continue;
        for (uint32_t f = 0; f < cnt; f++) {
            double q[3], d[3];
            const int ft = mesh_closest(p, &tile[9 * f], &tile[9 * f + 3], &tile[9 * f + 6], q);
            sub3v(p, q, d);
            const double d2 = dot3v(d, d);
            const uint32_t o = tile_orig[f];
            if (d2 < best || (d2 == best && o < bt)) {
                best = d2;
                bt = o;
                bf = ft;
                bq[0] = q[0], bq[1] = q[1], bq[2] = q[2];
            }
        }
    }
    if (!in) return;
    if (!node || bf < 0) {
        values[k] = 0.0;
        return;
    }
    double diff[3];
    sub3v(p, bq, diff);
    const double dist = sqrt(dot3v(diff, diff));
    const double *n = bf < 3 ? M.vn + 3 * (size_t)M.tv[3 * (size_t)bt + bf]
                     : (bf < 6 ? M.en + 3 * (size_t)M.te[3 * (size_t)bt + (bf - 3)]
                               : M.fn + 3 * (size_t)bt);
    values[k] = dot3v(diff, n) < 0.0 ? -dist : dist;
}

// CubicSerendipityDiscretization::new pruning (discretization.rs:208-236): a
// cell is dropped when all 32 node values lie below -prune or all above
// +prune; the nodes of kept cells are present
__global__ void k_sdf_cells(VmGrid g, const double *values, double prune, uint8_t *present) {
    const size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t ncell = (size_t)g.n[0] * g.n[1] * g.n[2];
    if (c >= ncell) return;
    const size_t nx = 3 * (size_t)g.n[0] + 1, ny = 3 * (size_t)g.n[1] + 1;
    const size_t cx = c % g.n[0], cy = (c / g.n[0]) % g.n[1], cz = c / ((size_t)g.n[0] * g.n[1]);
    const size_t base = (3 * cz * ny + 3 * cy) * nx + 3 * cx;
    bool low = true, high = true;
    for (int i = 0; i < 32; i++) {
        const double v =
            values[base + ((size_t)kNodeOff[i][2] * ny + kNodeOff[i][1]) * nx + kNodeOff[i][0]];
        low = low && v < -prune;
        high = high && v > prune;
    }
    if (low || high) return;
    for (int i = 0; i < 32; i++)
        present[base + ((size_t)kNodeOff[i][2] * ny + kNodeOff[i][1]) * nx + kNodeOff[i][0]] = 1;
}

}  // namespace sph

