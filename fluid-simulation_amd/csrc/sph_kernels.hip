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
// block reductions (deterministic fixed-shape trees)
// ------------------------------------------------------------------------
template <int N>
__device__ __forceinline__ void block_sum(double (&v)[N], double *lds /* N*kBlock */) {
    const int t = threadIdx.x;
#pragma unroll
    for (int k = 0; k < N; k++) lds[k * kBlock + t] = v[k];
    __syncthreads();
    for (int s = kBlock / 2; s > 0; s >>= 1) {
        if (t < s) {
#pragma unroll
            for (int k = 0; k < N; k++) lds[k * kBlock + t] += lds[k * kBlock + t + s];
        }
        __syncthreads();
    }
#pragma unroll
    for (int k = 0; k < N; k++) v[k] = lds[k * kBlock];
    __syncthreads();
}

__global__ void k_fill_u32(uint32_t *p, size_t n, uint32_t v) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_fill_f64(double *p, size_t n, double v) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// sum of `rows` rows of `cols` partials each: out[r] (+)= sum_c in[c*rows + r]
__global__ void k_reduce_rows(const double *in, int cols, int rows, double *out, int accumulate) {
    __shared__ double lds[kBlock];
    const int r = blockIdx.x;
    double s = 0.0;
    for (int c = threadIdx.x; c < cols; c += kBlock) s += in[(size_t)c * rows + r];
    double v[1] = {s};
    block_sum<1>(v, lds);
    if (threadIdx.x == 0) out[r] = accumulate ? out[r] + v[0] : v[0];
}

// ------------------------------------------------------------------------
// integration (integration_schemes/euler_cromer.rs:12-25)
// ------------------------------------------------------------------------
__global__ void k_integrate_ec(size_t n, double dt, double *x, double *y, double *z, double *vx,
                               double *vy, double *vz, const double *ax, const double *ay,
                               const double *az) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double nvx = vx[i] + dt * ax[i], nvy = vy[i] + dt * ay[i], nvz = vz[i] + dt * az[i];
    vx[i] = nvx;
    vy[i] = nvy;
    vz[i] = nvz;
    x[i] = x[i] + dt * nvx;
    y[i] = y[i] + dt * nvy;
    z[i] = z[i] + dt * nvz;
}

// ------------------------------------------------------------------------
// pre-update scan: cell coordinates, occupied-cell bounding box, max speed,
// count of particles below the disable threshold (sph/mod.rs:677-691)
// ------------------------------------------------------------------------
__device__ __forceinline__ int sat_i32(double v) {
    if (v != v) return 0;
    if (v >= 2147483647.0) return 2147483647;
    if (v <= -2147483648.0) return (int)(-2147483647 - 1);
    return (int)v;
}

// per-block partials of the occupied-cell box (min x,y,z; max x,y,z); a
// single-block pass folds them (no contended atomics on six words)
__global__ void k_cells(size_t n, double cs, const double *x, const double *y, const double *z,
                        int *ci, int *bbox_partials) {
    __shared__ int lds[6][kBlock / 64];
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    int v[6] = {2147483647, 2147483647, 2147483647, -2147483647 - 1, -2147483647 - 1,
                -2147483647 - 1};
    if (i < n) {
        int a = sat_i32(floor(x[i] / cs)), b = sat_i32(floor(y[i] / cs)),
            c = sat_i32(floor(z[i] / cs));
        ci[3 * i] = a;
        ci[3 * i + 1] = b;
        ci[3 * i + 2] = c;
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

// per-block max |v| (calc_max_speed, sph/mod.rs:424-441) and count of z < threshold
__global__ void k_speed_below(size_t n, const double *vx, const double *vy, const double *vz,
                              const double *z, double thr, double *partials /* 2 per block */) {
    __shared__ double lds[2][kBlock / 64];
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    double s = 0.0, b = 0.0;
    if (i < n) {
        s = sqrt(dot3(vx[i], vy[i], vz[i], vx[i], vy[i], vz[i]));
        b = z[i] < thr ? 1.0 : 0.0;
    }
    for (int off = 32; off > 0; off >>= 1) {
        s = fmax(s, __shfl_xor(s, off));
        b += __shfl_xor(b, off);
    }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        lds[0][w] = s;
        lds[1][w] = b;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int q = 1; q < kBlock / 64; q++) {
            s = fmax(s, lds[0][q]);
            b += lds[1][q];
        }
        partials[(size_t)blockIdx.x * 2] = s;
        partials[(size_t)blockIdx.x * 2 + 1] = b;
    }
}

__global__ void k_fold_speed(const double *partials, int nparts, double *out) {
    __shared__ double lds[2][kBlock];
    double s = 0.0, b = 0.0;
    for (int q = threadIdx.x; q < nparts; q += kBlock) {
        s = fmax(s, partials[(size_t)q * 2]);
        b += partials[(size_t)q * 2 + 1];
    }
    lds[0][threadIdx.x] = s;
    lds[1][threadIdx.x] = b;
    __syncthreads();
    for (int t = kBlock / 2; t > 0; t >>= 1) {
        if (threadIdx.x < t) {
            lds[0][threadIdx.x] = fmax(lds[0][threadIdx.x], lds[0][threadIdx.x + t]);
            lds[1][threadIdx.x] += lds[1][threadIdx.x + t];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = lds[0][0];
        out[1] = lds[1][0];
    }
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
template <class T>
__global__ void k_scatter_back(size_t n, const uint32_t *perm, const T *src, T *dst) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[perm[i]] = src[i];
}
template __global__ void k_gather<double>(size_t, const uint32_t *, const double *, double *);
template __global__ void k_gather<uint32_t>(size_t, const uint32_t *, const uint32_t *,
                                             uint32_t *);
template __global__ void k_gather<int>(size_t, const uint32_t *, const int *, int *);
template __global__ void k_scatter_back<double>(size_t, const uint32_t *, const double *,
                                                double *);
template __global__ void k_scatter_back<uint8_t>(size_t, const uint32_t *, const uint8_t *,
                                                 uint8_t *);

// ------------------------------------------------------------------------
// neighbor search (spatial_hashing.rs:75-103): for every query, the samples
// of the (2c+1)^3 stencil with |x_q - x_s| < range. The stencil is walked in
// the reference's (dx, dy, dz) order; with z the fastest key axis each
// (dx, dy) row is one contiguous range of the sorted samples, and members of
// a cell are ordered by reference index, so the emitted order equals the
// reference's. The predicate d2 < dthr, with dthr the least double whose
// correctly rounded sqrt is >= range, is exactly sqrt(d2) < range.
// ------------------------------------------------------------------------
template <bool FILL>
__global__ void k_neighbors(size_t nq, const double *qx, const double *qy, const double *qz,
                            double cs, int reach, double dthr, GridDev g, const double *sx,
                            const double *sy, const double *sz, const uint32_t *off,
                            uint32_t *out_count, uint32_t *out_idx) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nq) return;
    const double px = qx[i], py = qy[i], pz = qz[i];
    const int cx = sat_i32(floor(px / cs)), cy = sat_i32(floor(py / cs)),
              cz = sat_i32(floor(pz / cs));
    uint32_t cnt = 0;
    uint32_t o = FILL ? off[i] : 0;
    int zlo = cz - reach - g.org[2], zhi = cz + reach - g.org[2];
    if (zlo < 0) zlo = 0;
    if (zhi > g.dim[2] - 1) zhi = g.dim[2] - 1;
    if (zlo <= zhi) {
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
                        if (FILL) out_idx[o + cnt] = j;
                        cnt++;
                    }
                }
            }
        }
    }
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

// ------------------------------------------------------------------------
// wavefront-sliced ELL pair cache. For slice s (64 consecutive sorted
// particles) the k-th neighbor of lane l lives at 64 (slice_off[s] + k) + l, so
// every load of a neighbor pass is one contiguous 64-lane access. Each slot
// holds the neighbor index and grad W_ij, computed once per update and reused
// by every pass of the step (the reference recomputes grad W in every pass).
// ------------------------------------------------------------------------
__global__ void k_count_ell(size_t nq, const double *qx, const double *qy, const double *qz,
                            double cs, int reach, double dthr, GridDev g, uint32_t *count,
                            uint32_t *slice_len) {
    const unsigned lb = xcd_block();
    size_t i = (size_t)lb * blockDim.x + threadIdx.x;
    uint32_t cnt = 0;
    if (i < nq) {
        const double px = qx[i], py = qy[i], pz = qz[i];
        const int cx = sat_i32(floor(px / cs)), cy = sat_i32(floor(py / cs)),
                  cz = sat_i32(floor(pz / cs));
        int zlo = cz - reach - g.org[2], zhi = cz + reach - g.org[2];
        if (zlo < 0) zlo = 0;
        if (zhi > g.dim[2] - 1) zhi = g.dim[2] - 1;
        if (zlo <= zhi) {
            for (int dx = -reach; dx <= reach; dx++) {
                int gx = cx + dx - g.org[0];
                if (gx < 0 || gx >= g.dim[0]) continue;
                for (int dy = -reach; dy <= reach; dy++) {
                    int gy = cy + dy - g.org[1];
                    if (gy < 0 || gy >= g.dim[1]) continue;
                    uint32_t row = ((uint32_t)gx * (uint32_t)g.dim[1] + (uint32_t)gy) * (uint32_t)g.dim[2];
                    uint32_t a = g.start[row + zlo], b = g.start[row + zhi + 1];
                    for (uint32_t j = a; j < b; j++) {
                        double ddx = px - qx[j], ddy = py - qy[j], ddz = pz - qz[j];
                        double d2 = (ddx * ddx + ddy * ddy) + ddz * ddz;
                        cnt += d2 < dthr ? 1u : 0u;
                    }
                }
            }
        }
        count[i] = cnt;
    }
    uint32_t mx = cnt;
    for (int off = 32; off > 0; off >>= 1) mx = max(mx, (uint32_t)__shfl_xor((int)mx, off));
    if ((threadIdx.x & 63) == 0 && i < nq) slice_len[i >> 6] = mx;
}

// neighbor indices into the ELL cache (reference order within each particle)
__global__ void k_fill_ell(size_t n, double cs, int reach, double dthr, GridDev g,
                           const double *x, const double *y, const double *z,
                           const uint32_t *slice_off, uint32_t *nj) {
    const unsigned lb = xcd_block();
    size_t i = (size_t)lb * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double px = x[i], py = y[i], pz = z[i];
    const int cx = sat_i32(floor(px / cs)), cy = sat_i32(floor(py / cs)),
              cz = sat_i32(floor(pz / cs));
    size_t e = (size_t)slice_off[i >> 6] * 64 + (i & 63);
    int zlo = cz - reach - g.org[2], zhi = cz + reach - g.org[2];
    if (zlo < 0) zlo = 0;
    if (zhi > g.dim[2] - 1) zhi = g.dim[2] - 1;
    if (zlo > zhi) return;
    for (int dx = -reach; dx <= reach; dx++) {
        int gx = cx + dx - g.org[0];
        if (gx < 0 || gx >= g.dim[0]) continue;
        for (int dy = -reach; dy <= reach; dy++) {
            int gy = cy + dy - g.org[1];
            if (gy < 0 || gy >= g.dim[1]) continue;
            uint32_t row = ((uint32_t)gx * (uint32_t)g.dim[1] + (uint32_t)gy) * (uint32_t)g.dim[2];
            uint32_t a = g.start[row + zlo], b = g.start[row + zhi + 1];
            for (uint32_t j = a; j < b; j++) {
                double rx = px - x[j], ry = py - y[j], rz = pz - z[j];
                double d2 = (rx * rx + ry * ry) + rz * rz;
                if (d2 < dthr) {
                    nj[e] = j;
                    e += 64;
                }
            }
        }
    }
}

// grad W per cached pair and get_volume (quantities.rs:14-61); the lanes of a
// wavefront step through the slots in lockstep, so every cache write is one
// contiguous 64-lane store
__global__ void k_pairs_volume(size_t n, KernelConst K, double V0, const double *x,
                               const double *y, const double *z, const uint32_t *cnt,
                               const uint32_t *slice_off, const uint32_t *nj, double *gxo,
                               double *gyo, double *gzo, const BodyDev *bodies, int nb,
                               double *V) {
    const unsigned lb = xcd_block();
    size_t i = (size_t)lb * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double px = x[i], py = y[i], pz = z[i];
    const uint32_t c = cnt[i];
    size_t e = (size_t)slice_off[i >> 6] * 64 + (i & 63);
    double accu = 0.;
    #pragma unroll 4
        for (uint32_t k = 0; k < c; k++, e += 64) {
        uint32_t j = nj[e];
        double rx = px - x[j], ry = py - y[j], rz = pz - z[j];
        double rn = sqrt((rx * rx + ry * ry) + rz * rz);
        double q = rn / K.h;
        accu += V0 * (K.wfac * kernel_w(q));
        double wx = 0.0, wy = 0.0, wz = 0.0;
        if (rn != 0.0) {
            double d = kernel_dqw(q);
            wx = ((K.gfac * rx) / rn) * d;
            wy = ((K.gfac * ry) / rn) * d;
            wz = ((K.gfac * rz) / rn) * d;
        }
        gxo[e] = wx;
        gyo[e] = wy;
        gzo[e] = wz;
    }
    for (int bi = 0; bi < nb; bi++) {
        const BodyDev &b = bodies[bi];
        if (b.kind == 0) {
            for (uint32_t k = b.soff[i]; k < b.soff[i + 1]; k++) {
                uint32_t s = b.sidx[k];
                accu += b.svol[s] * kernel_fn(K, px - b.sx[s], py - b.sy[s], pz - b.sz[s]);
            }
        } else if (b.flag[i]) {
            accu += b.bvol[i] * kernel_fn(K, px - b.bx[i], py - b.by[i], pz - b.bz[i]);
        }
    }
    V[i] = V0 / accu;
}

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

// ------------------------------------------------------------------------
// VolumeMapBoundary::find_boundary_samples, one body
// (volume_map_boundary.rs:1259-1424)
// ------------------------------------------------------------------------
__device__ inline int body_sd(const BodyDev &b, double px, double py, double pz, double &sd) {
    if (!b.dynamic) return serendipity(b.sdf, px, py, pz, sd);
    double lx, ly, lz;
    world_to_local(b, px, py, pz, lx, ly, lz);
    return serendipity(b.sdf, lx, ly, lz, sd);
}
__device__ inline int body_sdg(const BodyDev &b, double px, double py, double pz, double &gx,
                               double &gy, double &gz) {
    if (!b.dynamic) return serendipity_grad(b.sdf, px, py, pz, gx, gy, gz);
    double lx, ly, lz, ux, uy, uz;
    world_to_local(b, px, py, pz, lx, ly, lz);
    int e = serendipity_grad(b.sdf, lx, ly, lz, ux, uy, uz);
    if (e) return e;
    quat_rotate(b.q[0], b.q[1], b.q[2], b.q[3], ux, uy, uz, gx, gy, gz);
    return 0;
}
__device__ inline int body_vm(const BodyDev &b, double px, double py, double pz, double &v) {
    if (!b.dynamic) return serendipity(b.vm, px, py, pz, v);
    double lx, ly, lz;
    world_to_local(b, px, py, pz, lx, ly, lz);
    return serendipity(b.vm, lx, ly, lz, v);
}

__global__ void k_vm_sample(size_t n, const double *x, const double *y, const double *z,
                            BodyDev b, double range, double dxr, double vmax, uint8_t *flag,
                            double *bx, double *by, double *bz, double *bvx, double *bvy,
                            double *bvz, double *bvol) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double px = x[i], py = y[i], pz = z[i];
    uint8_t ok = 0;
    double sd, gx, gy, gz, vol;
    if (!body_sd(b, px, py, pz, sd) && sd < range && !body_sdg(b, px, py, pz, gx, gy, gz)) {
        double gn = sqrt(dot3(gx, gy, gz, gx, gy, gz));
        if (gn > 0.5 && gn < 2.0 && !body_vm(b, px, py, pz, vol) && vol > 0.) {
            vol = fmin(vol, vmax);
            double d = fmax(sd + 0.25 * dxr, dxr);
            double ox = px - (gx / gn) * d, oy = py - (gy / gn) * d, oz = pz - (gz / gn) * d;
            double vx, vy, vz;
            body_velocity(b, ox, oy, oz, vx, vy, vz);
            bx[i] = ox;
            by[i] = oy;
            bz[i] = oz;
            bvx[i] = vx;
            bvy[i] = vy;
            bvz[i] = vz;
            bvol[i] = vol;
            ok = 1;
        }
    }
    flag[i] = ok;
}

// the four viscosity samples around the normal sample (volume_map_boundary.rs:1347-1376)
__device__ inline void visc_samples(double px, double py, double pz, double ox, double oy,
                                    double oz, double range, double (&vp)[4][3]) {
    double vx = ox - px, vy = oy - py, vz = oz - pz;
    double vn = sqrt(dot3(vx, vy, vz, vx, vy, vz));
    double ux = vx / vn, uy = vy / vn, uz = vz / vn;
    double tx = 1., ty = 0., tz = 0.;
    if (fabs(dot3(ux, uy, uz, tx, ty, tz)) > 0.9) {
        tx = 0.;
        ty = 1.;
    }
    double t1x, t1y, t1z, t2x, t2y, t2z;
    cross3(vx, vy, vz, tx, ty, tz, t1x, t1y, t1z);
    cross3(vx, vy, vz, t1x, t1y, t1z, t2x, t2y, t2z);
    double n1 = sqrt(dot3(t1x, t1y, t1z, t1x, t1y, t1z));
    double n2 = sqrt(dot3(t2x, t2y, t2z, t2x, t2y, t2z));
    t1x = t1x / n1;
    t1y = t1y / n1;
    t1z = t1z / n1;
    t2x = t2x / n2;
    t2y = t2y / n2;
    t2z = t2z / n2;
    double d = 0.5 * range;
    vp[0][0] = ox + d * t1x; vp[0][1] = oy + d * t1y; vp[0][2] = oz + d * t1z;
    vp[1][0] = ox - d * t1x; vp[1][1] = oy - d * t1y; vp[1][2] = oz - d * t1z;
    vp[2][0] = ox + d * t2x; vp[2][1] = oy + d * t2y; vp[2][2] = oz + d * t2z;
    vp[3][0] = ox - d * t2x; vp[3][1] = oy - d * t2y; vp[3][2] = oz - d * t2z;
}

// ------------------------------------------------------------------------
// reset + gravity + artificial viscosity (non_pressure_accelerations.rs:19-185)
// with the reaction on dynamic bodies reduced per block in FP64.
// ------------------------------------------------------------------------
__global__ void k_nonpressure(size_t n, KernelConst K, NonPressureConst P, const double *x,
                              const double *y, const double *z, const double *vx,
                              const double *vy, const double *vz, const double *m,
                              const double *V, const uint32_t *cnt, const uint32_t *slice_off,
                              const uint32_t *nj, const double *gxa, const double *gya,
                              const double *gza, const BodyDev *bodies, int nb, double *ax,
                              double *ay, double *az, double *body_partials) {
    __shared__ double lds[6 * kBlock];
    const unsigned lb = xcd_block();
    size_t i = (size_t)lb * blockDim.x + threadIdx.x;
    const bool live = i < n;
    double px = 0, py = 0, pz = 0, ux = 0, uy = 0, uz = 0, mi = 0;
    double gx = 0.0, gy = 0.0, gz = 0.0;
    double acx = 0., acy = 0., acz = 0.;
    if (live) {
        px = x[i]; py = y[i]; pz = z[i];
        ux = vx[i]; uy = vy[i]; uz = vz[i];
        mi = m[i];
        if (P.gravity_mode == 0) {
            gx = 0.0 + 0.0;
            gy = 0.0 + 0.0;
            gz = 0.0 + -9.81;
        } else {
            double dx = P.gcenter[0] - px, dy = P.gcenter[1] - py, dz = P.gcenter[2] - pz;
            double dist = sqrt(dot3(dx, dy, dz, dx, dy, dz) + 1e-4);
            gx = 0.0 + (9.81 * dx) / dist;
            gy = 0.0 + (9.81 * dy) / dist;
            gz = 0.0 + (9.81 * dz) / dist;
        }
        const uint32_t c = cnt[i];
        size_t e = (size_t)slice_off[i >> 6] * 64 + (i & 63);
        #pragma unroll 4
        for (uint32_t k = 0; k < c; k++, e += 64) {
            uint32_t j = nj[e];
            double rx = px - x[j], ry = py - y[j], rz = pz - z[j];
            double wx = gxa[e], wy = gya[e], wz = gza[e];
            double cf = P.nu_f * 2. * (3. + 2.) * V[j] *
                        dot3(ux - vx[j], uy - vy[j], uz - vz[j], rx, ry, rz) /
                        (dot3(rx, ry, rz, rx, ry, rz) + P.eps2);
            acx += cf * wx;
            acy += cf * wy;
            acz += cf * wz;
        }
    }
    for (int bi = 0; bi < nb; bi++) {
        const BodyDev &b = bodies[bi];
        double f[6] = {0., 0., 0., 0., 0., 0.};
        if (live) {
            if (b.kind == 0) {
                for (uint32_t e = b.soff[i]; e < b.soff[i + 1]; e++) {
                    uint32_t s = b.sidx[e];
                    double rx = px - b.sx[s], ry = py - b.sy[s], rz = pz - b.sz[s];
                    double wx, wy, wz;
                    kernel_grad(K, rx, ry, rz, wx, wy, wz);
                    // static samples have zero velocity
                    double c = P.nu_b * 2. * (3. + 2.) * b.svol[s] *
                               dot3(ux - 0.0, uy - 0.0, uz - 0.0, rx, ry, rz) /
                               (dot3(rx, ry, rz, rx, ry, rz) + P.eps2);
                    acx += c * wx;
                    acy += c * wy;
                    acz += c * wz;
                }
            } else if (b.flag[i]) {
                double vp[4][3];
                visc_samples(px, py, pz, b.bx[i], b.by[i], b.bz[i], K.h, vp);
                double vv = 0.25 * b.bvol[i];
                for (int k = 0; k < 4; k++) {
                    double sxk = vp[k][0], syk = vp[k][1], szk = vp[k][2];
                    double svx, svy, svz;
                    body_velocity(b, sxk, syk, szk, svx, svy, svz);
                    double rx = px - sxk, ry = py - syk, rz = pz - szk;
                    double wx, wy, wz;
                    kernel_grad(K, rx, ry, rz, wx, wy, wz);
                    double c = P.nu_b * 2. * (3. + 2.) * vv *
                               dot3(ux - svx, uy - svy, uz - svz, rx, ry, rz) /
                               (dot3(rx, ry, rz, rx, ry, rz) + P.eps2);
                    double a0 = c * wx, a1 = c * wy, a2 = c * wz;
                    if (b.dynamic) {
                        double fx = -(mi * a0), fy = -(mi * a1), fz = -(mi * a2);
                        double tx, ty, tz;
                        cross3(sxk - b.com[0], syk - b.com[1], szk - b.com[2], fx, fy, fz, tx,
                               ty, tz);
                        f[0] += fx; f[1] += fy; f[2] += fz;
                        f[3] += tx; f[4] += ty; f[5] += tz;
                    }
                    acx += a0;
                    acy += a1;
                    acz += a2;
                }
            }
        }
        if (b.dynamic) {  // uniform across the block
            block_sum<6>(f, lds);
            if (threadIdx.x == 0) {
                double *dst = body_partials + ((size_t)lb * kMaxBodies + bi) * 6;
                for (int k = 0; k < 6; k++) dst[k] = f[k];
            }
        }
    }
    if (live) {
        ax[i] = gx + acx;
        ay[i] = gy + acy;
        az[i] = gz + acz;
    }
}

// ------------------------------------------------------------------------
// SESPH state equation (sesph.rs:40-51)
// ------------------------------------------------------------------------
__global__ void k_sesph_pressure(size_t n, double k, double V0, const double *V, double *p) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = k * fmax(V0 / V[i] - 1., 0.);
}

// ------------------------------------------------------------------------
// add_pressure_acceleration (pressure_solver/mod.rs:121-195)
// ------------------------------------------------------------------------
__global__ void k_pressure_accel(size_t n, KernelConst K, double gamma, const double *x,
                                 const double *y, const double *z, const double *m,
                                 const double *V, const double *p, const uint32_t *cnt,
                                 const uint32_t *slice_off, const uint32_t *nj,
                                 const double *gxa, const double *gya, const double *gza,
                                 const BodyDev *bodies, int nb, int overwrite, int mirror,
                                 double *tx_, double *ty_, double *tz_, double *body_partials) {
    __shared__ double lds[6 * kBlock];
    const unsigned lb = xcd_block();
    size_t i = (size_t)lb * blockDim.x + threadIdx.x;
    const bool live = i < n;
    double px = 0, py = 0, pz = 0, Vi = 0, mi = 1, pi = 0;
    double acx = 0., acy = 0., acz = 0.;
    if (live) {
        Vi = V[i]; mi = m[i]; pi = p[i];
        const uint32_t c = cnt[i];
        size_t e = (size_t)slice_off[i >> 6] * 64 + (i & 63);
        const double vm = Vi / mi;
        #pragma unroll 4
        for (uint32_t k = 0; k < c; k++, e += 64) {
            uint32_t j = nj[e];
            double cf = vm * V[j] * (pi + p[j]);
            acx -= cf * gxa[e];
            acy -= cf * gya[e];
            acz -= cf * gza[e];
        }
        if (nb) {
            px = x[i]; py = y[i]; pz = z[i];
        }
    }
    for (int bi = 0; bi < nb; bi++) {
        const BodyDev &b = bodies[bi];
        double f[6] = {0., 0., 0., 0., 0., 0.};
        if (live) {
            if (b.kind == 0) {
                for (uint32_t e = b.soff[i]; e < b.soff[i + 1]; e++) {
                    uint32_t s = b.sidx[e];
                    double wx, wy, wz;
                    kernel_grad(K, px - b.sx[s], py - b.sy[s], pz - b.sz[s], wx, wy, wz);
                    double c = 2. * gamma * Vi * b.svol[s] * pi;
                    acx -= (c * wx) / mi;
                    acy -= (c * wy) / mi;
                    acz -= (c * wz) / mi;
                }
            } else if (b.flag[i]) {
                double sxk = b.bx[i], syk = b.by[i], szk = b.bz[i];
                double wx, wy, wz;
                kernel_grad(K, px - sxk, py - syk, pz - szk, wx, wy, wz);
                double c = 2. * gamma * Vi * b.bvol[i] * pi;
                double fx = c * wx, fy = c * wy, fz = c * wz;
                if (mirror && b.dynamic) {
                    double tx, ty, tz;
                    cross3(sxk - b.com[0], syk - b.com[1], szk - b.com[2], fx, fy, fz, tx, ty,
                           tz);
                    f[0] += fx; f[1] += fy; f[2] += fz;
                    f[3] += tx; f[4] += ty; f[5] += tz;
                }
                acx -= fx / mi;
                acy -= fy / mi;
                acz -= fz / mi;
            }
        }
        if (mirror && b.dynamic) {
            block_sum<6>(f, lds);
            if (threadIdx.x == 0) {
                double *dst = body_partials + ((size_t)lb * kMaxBodies + bi) * 6;
                for (int k = 0; k < 6; k++) dst[k] = f[k];
            }
        }
    }
    if (live) {
        if (overwrite) {
            tx_[i] = acx;
            ty_[i] = acy;
            tz_[i] = acz;
        } else {
            tx_[i] += acx;
            ty_[i] += acy;
            tz_[i] += acz;
        }
    }
}

// ------------------------------------------------------------------------
// IISPH pieces (iisph.rs:117-376)
// ------------------------------------------------------------------------
// v* = (v or v*) + dt a   (pressure_solver/mod.rs:90-108)
__global__ void k_pred_vel(size_t n, double dt, int to_pred, const double *vx, const double *vy,
                           const double *vz, const double *ax, const double *ay,
                           const double *az, double *svx, double *svy, double *svz) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double bx = to_pred ? svx[i] : vx[i], by = to_pred ? svy[i] : vy[i],
           bz = to_pred ? svz[i] : vz[i];
    svx[i] = bx + dt * ax[i];
    svy[i] = by + dt * ay[i];
    svz[i] = bz + dt * az[i];
}

// source term (vp or vde), diagonal element and initial pressure in one
// neighbor pass (set_source_term_* :117-251, set_diagonal_element :254-342,
// initialize :344-376 with clamp_pressure = true)
__global__ void k_source_diag(size_t n, KernelConst K, SolveConst S, int vp, const double *x,
                              const double *y, const double *z, const double *svx,
                              const double *svy, const double *svz, const double *m,
                              const double *V, const uint32_t *cnt, const uint32_t *slice_off,
                              const uint32_t *nj, const double *gxa, const double *gya,
                              const double *gza, const BodyDev *bodies, int nb, double *s_f,
                              double *a_ff, double *p) {
    const unsigned lb = xcd_block();
    size_t i = (size_t)lb * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double px = x[i], py = y[i], pz = z[i];
    const double ux = svx[i], uy = svy[i], uz = svz[i];
    const double Vi = V[i];
    const double dt = S.dt, dt2 = S.dt * S.dt;
    double src = vp ? 1. - S.V0 / Vi : 0.;
    double sfx = 0., sfy = 0., sfz = 0., sf2 = 0., sbx = 0., sby = 0., sbz = 0.;
    const uint32_t c = cnt[i];
    size_t e = (size_t)slice_off[i >> 6] * 64 + (i & 63);
    #pragma unroll 4
        for (uint32_t k = 0; k < c; k++, e += 64) {
        uint32_t j = nj[e];
        double wx = gxa[e], wy = gya[e], wz = gza[e];
        double Vj = V[j];
        src -= dt * Vj * dot3(ux - svx[j], uy - svy[j], uz - svz[j], wx, wy, wz);
        sfx += Vj * wx;
        sfy += Vj * wy;
        sfz += Vj * wz;
        sf2 -= dt2 * Vi * powi2(Vj) / m[j] * dot3(wx, wy, wz, wx, wy, wz);
    }
    for (int bi = 0; bi < nb; bi++) {
        const BodyDev &b = bodies[bi];
        if (b.kind == 0) {
            for (uint32_t e = b.soff[i]; e < b.soff[i + 1]; e++) {
                uint32_t s = b.sidx[e];
                double wx, wy, wz;
                kernel_grad(K, px - b.sx[s], py - b.sy[s], pz - b.sz[s], wx, wy, wz);
                double Vb = b.svol[s];
                src -= dt * Vb * dot3(ux - 0.0, uy - 0.0, uz - 0.0, wx, wy, wz);
                sbx += Vb * wx;
                sby += Vb * wy;
                sbz += Vb * wz;
            }
        } else if (b.flag[i]) {
            double wx, wy, wz;
            kernel_grad(K, px - b.bx[i], py - b.by[i], pz - b.bz[i], wx, wy, wz);
            double Vb = b.bvol[i];
            src -= dt * Vb * dot3(ux - b.bvx[i], uy - b.bvy[i], uz - b.bvz[i], wx, wy, wz);
            sbx += Vb * wx;
            sby += Vb * wy;
            sbz += Vb * wz;
        }
    }
    double ci = -Vi / m[i];
    double cfx = ci * (sfx + 2. * S.gamma * sbx), cfy = ci * (sfy + 2. * S.gamma * sby),
           cfz = ci * (sfz + 2. * S.gamma * sbz);
    double aii = dt2 * dot3(cfx, cfy, cfz, sfx + sbx, sfy + sby, sfz + sbz) + sf2;
    s_f[i] = src;
    a_ff[i] = aii;
    p[i] = (aii > S.eps || aii < -S.eps) ? fmax(S.omega * src / aii, 0.) : 0.;
}

// relaxed Jacobi update and per-block sum of the predicted density error
// (iisph.rs:440-531)
__global__ void k_jacobi(size_t n, KernelConst K, SolveConst S, const double *x, const double *y,
                         const double *z, const double *V, const double *pax, const double *pay,
                         const double *paz, const double *s_f, const double *a_ff,
                         const uint32_t *cnt, const uint32_t *slice_off, const uint32_t *nj,
                         const double *gxa, const double *gya, const double *gza,
                         const BodyDev *bodies, int nb, double *p, double *err_partials) {
    __shared__ double lds[kBlock];
    const unsigned lb = xcd_block();
    size_t i = (size_t)lb * blockDim.x + threadIdx.x;
    double err = 0.;
    if (i < n) {
        const double ux = pax[i], uy = pay[i], uz = paz[i];
        const double dt2 = S.dt * S.dt;
        double adp = 0.;
        const uint32_t c = cnt[i];
        size_t e = (size_t)slice_off[i >> 6] * 64 + (i & 63);
        #pragma unroll 4
        for (uint32_t k = 0; k < c; k++, e += 64) {
            uint32_t j = nj[e];
            adp += dt2 * V[j] * dot3(ux - pax[j], uy - pay[j], uz - paz[j], gxa[e], gya[e], gza[e]);
        }
        double px = 0, py = 0, pz = 0;
        if (nb) {
            px = x[i]; py = y[i]; pz = z[i];
        }
        for (int bi = 0; bi < nb; bi++) {
            const BodyDev &b = bodies[bi];
            if (b.kind == 0) {
                for (uint32_t e = b.soff[i]; e < b.soff[i + 1]; e++) {
                    uint32_t s = b.sidx[e];
                    double wx, wy, wz;
                    kernel_grad(K, px - b.sx[s], py - b.sy[s], pz - b.sz[s], wx, wy, wz);
                    adp += dt2 * b.svol[s] * dot3(ux, uy, uz, wx, wy, wz);
                }
            } else if (b.flag[i]) {
                double wx, wy, wz;
                kernel_grad(K, px - b.bx[i], py - b.by[i], pz - b.bz[i], wx, wy, wz);
                adp += dt2 * b.bvol[i] * dot3(ux, uy, uz, wx, wy, wz);
            }
        }
        const double aii = a_ff[i], si = s_f[i];
        if (aii < -S.eps) {
            double pn = p[i] + S.omega * (si - adp) / aii;
            p[i] = fmax(pn, 0.);
        }
        err = si < 0. ? fabs(adp - si) : 0.;
    }
    double v[1] = {err};
    block_sum<1>(v, lds);
    if (threadIdx.x == 0) err_partials[lb] = v[0];
}

// IISPHwOST: x* = x + dt v*   (iisph_optimized_source_term.rs:103-113)
__global__ void k_pred_pos(size_t n, double dt, const double *x, const double *y,
                           const double *z, const double *svx, const double *svy,
                           const double *svz, double *sx, double *sy, double *sz) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    sx[i] = x[i] + dt * svx[i];
    sy[i] = y[i] + dt * svy[i];
    sz[i] = z[i] + dt * svz[i];
}

// IISPHwOST: x* += dt^2 a, v* = v* + dt a   (iisph_optimized_source_term.rs:136-203)
__global__ void k_ost_commit(size_t n, double dt, const double *ax, const double *ay,
                             const double *az, double *sx, double *sy, double *sz, double *svx,
                             double *svy, double *svz) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double dt2 = dt * dt;
    sx[i] = sx[i] + dt2 * ax[i];
    sy[i] = sy[i] + dt2 * ay[i];
    sz[i] = sz[i] + dt2 * az[i];
    svx[i] = svx[i] + dt * ax[i];
    svy[i] = svy[i] + dt * ay[i];
    svz[i] = svz[i] + dt * az[i];
}

// ------------------------------------------------------------------------
// measurement reductions (sph/mod.rs:444-564)
// ------------------------------------------------------------------------
__global__ void k_measure(size_t n, double V0, const double *V, const double *m, const double *vx,
                          const double *vy, const double *vz, double *partials /* 3 per block */) {
    __shared__ double lds[3 * kBlock];
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    double v[3] = {0., 0., 0.};
    if (i < n) {
        double Vi = V[i], mi = m[i];
        v[0] = Vi < V0 ? mi / Vi : mi / V0;
        v[1] = Vi < V0 ? V0 / Vi : 1.;
        v[2] = 0.5 * mi * dot3(vx[i], vy[i], vz[i], vx[i], vy[i], vz[i]);
    }
    block_sum<3>(v, lds);
    if (threadIdx.x == 0)
        for (int k = 0; k < 3; k++) partials[(size_t)blockIdx.x * 3 + k] = v[k];
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

__global__ void k_widen_u32(size_t n, const uint32_t *in, uint64_t *out) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = (uint64_t)in[i];
}

}  // namespace sph
