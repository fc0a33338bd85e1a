// Per-step passes of System::update as multi-scene launches (gfx950).
//
// Every pass evaluates, per particle, the reference's expression in the
// reference's neighbor order in binary64 without FMA contraction, so a scene's
// results do not depend on which other scenes share the launch.
#include <hip/hip_runtime.h>

#include "sph_batch.hpp"

namespace sph {

// launch bounds of the pair passes: WAVES waves per SIMD exactly (registers
// for WAVES waves, and no fewer registers to reach more), with 256-thread
// workgroups; WAVES = 0 keeps the compiler's default bounds
#define PAIR_PASS_BOUNDS                                                       \
    __attribute__((amdgpu_flat_work_group_size(WAVES > 0 ? kBlock : 1,         \
                                               WAVES > 0 ? kBlock : 1024),     \
                   amdgpu_waves_per_eu(WAVES > 0 ? WAVES : 1, WAVES > 0 ? WAVES : 8)))

// scene of workgroup lb: the largest s with blk[s] <= lb (block-uniform)
template <class P>
__device__ __forceinline__ int scene_of(P blk, int ns, uint32_t lb) {
    int lo = 0, hi = ns;
    while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if (blk[mid] <= lb) lo = mid;
        else hi = mid;
    }
    return lo;
}

// DESC: the launch reads B from device memory over a grid rounded up from
// the batch's B.pblk[B.ns] workgroups (the form a captured graph replays);
// workgroups past them return at once
// (the pointers of a B read from memory are marked global: the compiler can
// infer that only for kernel arguments)
#define PARTICLE_THREAD_G(B, DESC)                                         \
    bool ok_ = true;                                                       \
    const unsigned lb = DESC ? xcd_block_of(gm((B).pblk)[(B).ns], ok_) : xcd_block(); \
    if (!ok_) return;                                                      \
    const int sk_ = scene_of(gm((B).pblk), (B).ns, lb);                    \
    const int sid = (B).smap ? (int)gm((B).smap)[sk_] : sk_;               \
    const auto &v = gm((B).views)[sid];                                    \
    const size_t i = (size_t)(lb - gm((B).pblk)[sk_]) * kBlock + threadIdx.x;  \
    const bool live = i < v.n;                                             \
    const size_t gi = (size_t)v.part_base + i;                             \
    (void)gi;                                                              \
    (void)live

#define PARTICLE_THREAD(B)                                                 \
    const unsigned lb = xcd_block();                                       \
    const int sk_ = scene_of((B).pblk, (B).ns, lb);                        \
    const int sid = (B).smap ? (int)(B).smap[sk_] : sk_;                   \
    const SceneView &v = (B).views[sid];                                   \
    const size_t i = (size_t)(lb - (B).pblk[sk_]) * kBlock + threadIdx.x;  \
    const bool live = i < v.n;                                             \
    const size_t gi = (size_t)v.part_base + i;                             \
    (void)gi;                                                              \
    (void)live

// Pair-cache slots: a slice's neighbor rows are stored in pairs, so that
// neighbors 2r and 2r+1 of a lane are adjacent (one 8-byte load reads both):
// neighbor k of lane l of slice s sits at 64 slice_off[s] + 128 (k / 2) +
// 2 l + k % 2. Slices hold an even number of rows.
#define ELL_BASE(B, v, i) ((size_t)gm((B).slice_off)[(v).slice_base + ((i) >> 6)] * 64 + 2 * ((i)&63))
#define ELL_AT(e, k) ((e) + 128 * ((size_t)(k) >> 1) + ((k)&1))

// The field pointers of the scene and body tables are read from device
// memory, so the compiler cannot tell their address space and would emit FLAT
// accesses; gm() marks them as global memory (GLOBAL loads and stores, SGPR
// base addressing, no LDS-counter waits).
#define GLOBAL_AS __attribute__((address_space(1)))
template <class T>
__device__ __forceinline__ GLOBAL_AS T *gm(T *p) {
    return (GLOBAL_AS T *)p;
}

// a[q * stride] summed for q = q0, q0 + kBlock, ... < q1, added in that order
// (one thread's share of a fixed-shape fold); eight loads are issued before
// their additions, which keeps the order and so the bits
__device__ __forceinline__ double strided_sum(const double *a_, uint32_t q0, uint32_t q1,
                                              size_t stride) {
    const GLOBAL_AS double *a = gm(a_);
    double acc = 0.0;
    uint32_t q = q0;
    for (; q1 > 7u * kBlock && q < q1 - 7u * kBlock; q += 8u * kBlock) {
        double v[8];
#pragma unroll
        for (int u = 0; u < 8; u++) v[u] = a[(size_t)(q + u * kBlock) * stride];
#pragma unroll
        for (int u = 0; u < 8; u++) acc += v[u];
    }
    for (; q < q1; q += kBlock) acc += a[(size_t)q * stride];
    return acc;
}

// packed per-particle records gathered by the neighbor loops (one record
// instead of separate 8-byte loads): q4[i] = {x, y, z, V} (32 bytes),
// pa3[i] = {pressure acceleration x, y, z} and v3[i] = {velocity x, y, z}
// (24 bytes, unpadded: the gathers' cost grows with the bytes per neighbor)
typedef double rec4 __attribute__((ext_vector_type(4)));
typedef double rec2 __attribute__((ext_vector_type(2)));
typedef double rec2u __attribute__((ext_vector_type(2), aligned(8)));  // 8-byte aligned
struct rec3 {
    double x, y, z;
};
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ rec4 ld4(const double *a, size_t i) {
    return ((const GLOBAL_AS rec4 *)a)[i];
}
// Neighbor gathers address a record by a 32-bit byte offset from the
// workgroup-uniform array base, which the compiler issues with the base in
// SGPRs and one 32-bit offset per lane (no 64-bit address arithmetic); the
// host keeps every scene below 2^27 particles so that j << 5 fits.
__device__ __forceinline__ rec4 ldj4(const double *a, uint32_t j) {
    return *(const GLOBAL_AS rec4 *)((const GLOBAL_AS char *)a + (j << 5));
}
__device__ __forceinline__ double ldj(const double *a, uint32_t j) {
    return *(const GLOBAL_AS double *)((const GLOBAL_AS char *)a + (j << 3));
}
// 24-byte record j: one 16-byte and one 8-byte load (the host keeps every
// scene below 2^27 particles, so 24 j fits 32 bits)
__device__ __forceinline__ rec3 ldj3(const double *a, uint32_t j) {
    const GLOBAL_AS char *r = (const GLOBAL_AS char *)a + j * 24u;
    const rec2u xy = *(const GLOBAL_AS rec2u *)r;
    return rec3{xy.x, xy.y, *(const GLOBAL_AS double *)(r + 16)};
}
__device__ __forceinline__ rec3 ld3(const double *a, size_t i) {
    const GLOBAL_AS double *r = (const GLOBAL_AS double *)a + 3 * i;
    const rec2u xy = *(const GLOBAL_AS rec2u *)r;
    return rec3{xy.x, xy.y, r[2]};
}
__device__ __forceinline__ void st3(double *a, size_t i, double x, double y, double z) {
    GLOBAL_AS double *r = (GLOBAL_AS double *)a + 3 * i;
    *(GLOBAL_AS rec2u *)r = rec2u{x, y};
    r[2] = z;
}
__device__ __forceinline__ void st4(double *a, size_t i, double x, double y, double z,
                                    double w) {
    rec4 r = {x, y, z, w};
    ((GLOBAL_AS rec4 *)a)[i] = r;
}

// grad W of pair slot e: read from the pair cache (CACHED) or recomputed from
// r = x_i - x_j with the same expression (bit-identical either way)
template <bool CACHED, bool ZS = false>
__device__ __forceinline__ void pair_grad(const BatchDev &B, const KernelConst &K, double rh,
                                          size_t e, double rx, double ry, double rz, double &wx,
                                          double &wy, double &wz) {
    if (CACHED) {
        wx = gm(B.ell_gx)[e];
        wy = gm(B.ell_gy)[e];
        wz = gm(B.ell_gz)[e];
    } else {
        kernel_grad_rh<ZS>(K, rh, rx, ry, rz, wx, wy, wz);
    }
}

// The neighbor loop of a pair pass: body(k, r) for k = 0 .. c-1 in order,
// where r = load(j) gathers the records of slot k's neighbor j (ej points at
// the particle's first slot; slot k is ej[ELL_AT(0, k)]). Three forms, all
// calling body in the same order with the same records:
//   PIPE = 0: the plain loop (index, then records, then the arithmetic);
//   PIPE = 1: software-pipelined: slot k+1's records and slot k+2's index are
//     in flight while slot k is evaluated;
//   PIPE = 2: two neighbors per step: one 8-byte load reads both indices and
//     both records are gathered before either is evaluated (for an odd count
//     the second index slot is read but not used: the record load repeats
//     the first neighbor);
//   PIPE = 3: UNROLL neighbors per step while UNROLL slots remain, all
//     their records gathered before the first is evaluated (a scheduling
//     barrier keeps the compiler from sinking the gathers between the
//     evaluations to save registers), then the remaining slots one by one.
// Prefetch slots clamp to c-1 and a pair's second slot lies inside the
// slice's even row count, so every load stays in bounds.
template <int PIPE, int UNROLL, class Load, class Body>
__device__ __forceinline__ void neighbor_loop(const GLOBAL_AS uint32_t *ej, uint32_t c, Load load,
                                              Body body) {
    if (c == 0) return;
    if (PIPE == 1) {
        uint32_t j1 = ej[ELL_AT(0, min(1u, c - 1))];
        auto r0 = load(ej[0]);
#pragma unroll UNROLL
        for (uint32_t k = 0; k < c; k++) {
            const uint32_t j2 = ej[ELL_AT(0, min(k + 2, c - 1))];
            const auto r1 = load(j1);
            body(k, r0);
            j1 = j2;
            r0 = r1;
        }
    } else if (PIPE == 3) {
        uint32_t k = 0;
        for (; k + UNROLL <= c; k += UNROLL) {
            decltype(load(0u)) r[UNROLL];
#pragma unroll
            for (int u = 0; u < UNROLL; u++) r[u] = load(ej[ELL_AT(0, k + u)]);
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int u = 0; u < UNROLL; u++) body(k + u, r[u]);
        }
#pragma nounroll
        for (; k < c; k++) body(k, load(ej[ELL_AT(0, k)]));
    } else if (PIPE == 2) {
        for (uint32_t k = 0; k < c; k += 2) {
            const u32x2 jj = *(const GLOBAL_AS u32x2 *)(ej + 128 * (size_t)(k >> 1));
            const uint32_t j1 = k + 1 < c ? jj.y : jj.x;
            const auto r0 = load(jj.x);
            const auto r1 = load(j1);
            body(k, r0);
            if (k + 1 < c) body(k + 1, r1);
        }
    } else {
#pragma unroll UNROLL
        for (uint32_t k = 0; k < c; k++) body(k, load(ej[ELL_AT(0, k)]));
    }
}

// gathered records of one neighbor
struct QU {  // {x, y, z, V} and a second 32-byte record
    rec4 q, u;
};
struct QV {  // {x, y, z, V} and a 24-byte record
    rec4 q;
    rec3 u;
};
struct QP {  // {x, y, z, V} and the pressure
    rec4 q;
    double p;
};

// ------------------------------------------------------------------------
// integration (integration_schemes/euler_cromer.rs:12-25)
// ------------------------------------------------------------------------
// explicit_euler.rs:12-25 (the position moves with the old velocity) and
// verlet.rs:13-34 (x' = x + (x - x_prev) + dt^2 a, v = (x' - x) / dt, the slot
// takes x); TakePredicted swaps slots on the host
__global__ void kb_integrate_ec(BatchDev B) {
    PARTICLE_THREAD(B);
    if (!live) return;
    const double dt = v.S.dt;
    double *X[3] = {v.x, v.y, v.z}, *U[3] = {v.vx, v.vy, v.vz};
    const double *A[3] = {v.ax, v.ay, v.az};
    if (v.integrator == SPH_INTEGRATOR_EULER_CROMER) {
        for (int k = 0; k < 3; k++) {
            const double nv = gm(U[k])[i] + dt * gm(A[k])[i];
            gm(U[k])[i] = nv;
            gm(X[k])[i] = gm(X[k])[i] + dt * nv;
        }
    } else if (v.integrator == SPH_INTEGRATOR_EXPLICIT_EULER) {
        for (int k = 0; k < 3; k++) {
            const double x0 = gm(X[k])[i], v0 = gm(U[k])[i];
            gm(X[k])[i] = x0 + dt * v0;
            gm(U[k])[i] = v0 + dt * gm(A[k])[i];
        }
    } else if (v.integrator == SPH_INTEGRATOR_VERLET) {
        double *P[3] = {v.ix, v.iy, v.iz};
        for (int k = 0; k < 3; k++) {
            const double x0 = gm(X[k])[i];
            const double nx = x0 + (x0 - gm(P[k])[i]) + (dt * dt) * gm(A[k])[i];
            gm(U[k])[i] = (nx - x0) / dt;
            gm(P[k])[i] = x0;
            gm(X[k])[i] = nx;
        }
    }
}

// ------------------------------------------------------------------------
// max |v| (calc_max_speed, sph/mod.rs:424-441) and count below the disable
// threshold (sph/mod.rs:677-686), per workgroup
// ------------------------------------------------------------------------
__global__ void kb_speed_below(BatchDev B) {
    __shared__ double lds[2][kBlock / 64];
    PARTICLE_THREAD(B);
    double sp = 0.0, b = 0.0;
    if (live) {
        sp = sqrt(dot3(gm(v.vx)[i], gm(v.vy)[i], gm(v.vz)[i], gm(v.vx)[i], gm(v.vy)[i], gm(v.vz)[i]));
        b = gm(v.z)[i] < v.disable_below ? 1.0 : 0.0;
    }
    for (int off = 32; off > 0; off >>= 1) {
        sp = fmax(sp, __shfl_xor(sp, off));
        b += __shfl_xor(b, off);
    }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        lds[0][w] = sp;
        lds[1][w] = b;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int q = 1; q < kBlock / 64; q++) {
            sp = fmax(sp, lds[0][q]);
            b += lds[1][q];
        }
        gm(B.partials)[(size_t)lb * 2] = sp;
        gm(B.partials)[(size_t)lb * 2 + 1] = b;
    }
}

// ------------------------------------------------------------------------
// cell coordinates (spatial_hashing.rs:40-46) and per-workgroup cell box
// ------------------------------------------------------------------------
__global__ void kb_cells(BatchDev B) {
    __shared__ int lds[6][kBlock / 64];
    PARTICLE_THREAD(B);
    int c[6] = {2147483647, 2147483647, 2147483647, -2147483647 - 1, -2147483647 - 1,
                -2147483647 - 1};
    if (live) {
        int a = sat_i32(floor(gm(v.x)[i] / v.cs)), b = sat_i32(floor(gm(v.y)[i] / v.cs)),
            d = sat_i32(floor(gm(v.z)[i] / v.cs));
        gm(B.ci)[3 * gi] = a;
        gm(B.ci)[3 * gi + 1] = b;
        gm(B.ci)[3 * gi + 2] = d;
        c[0] = c[3] = a;
        c[1] = c[4] = b;
        c[2] = c[5] = d;
    }
    for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
        for (int k = 0; k < 3; k++) c[k] = min(c[k], __shfl_xor(c[k], off));
#pragma unroll
        for (int k = 3; k < 6; k++) c[k] = max(c[k], __shfl_xor(c[k], off));
    }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0)
        for (int k = 0; k < 6; k++) lds[k][w] = c[k];
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int q = 1; q < kBlock / 64; q++) {
            for (int k = 0; k < 3; k++) c[k] = min(c[k], lds[k][q]);
            for (int k = 3; k < 6; k++) c[k] = max(c[k], lds[k][q]);
        }
        int *dst = reinterpret_cast<int *>(B.partials) + (size_t)lb * 6;
        for (int k = 0; k < 6; k++) dst[k] = c[k];
    }
}

// ------------------------------------------------------------------------
// counting sort by (cell key, reference index); z is the fastest key axis
// ------------------------------------------------------------------------
__device__ __forceinline__ uint32_t local_key(const GridDev &g, int cx, int cy, int cz) {
    return ((uint32_t)(cx - g.org[0]) * (uint32_t)g.dim[1] + (uint32_t)(cy - g.org[1])) *
               (uint32_t)g.dim[2] +
           (uint32_t)(cz - g.org[2]);
}

// ctr[key] += (active lanes holding key), one atomic per distinct key of the
// wavefront; returns the counter's value before the group's addition plus
// the lane's rank among the group's lanes. Particles arrive in the previous
// step's cell order, so a wavefront holds few distinct cells.
__device__ __forceinline__ uint32_t wave_group_add(uint32_t *ctr, uint32_t key) {
    uint64_t todo = __ballot(1);
    const int lane = __lane_id();
    uint32_t mine = 0;
    while (todo) {
        const int leader = __ffsll((unsigned long long)todo) - 1;
        const uint32_t lk = __shfl(key, leader);
        const uint64_t grp = __ballot(key == lk);
        uint32_t old = 0;
        if (lane == leader) old = atomicAdd(&ctr[lk], (uint32_t)__popcll(grp));
        old = __shfl(old, leader);
        if (key == lk) mine = old + (uint32_t)__popcll(grp & __lanemask_lt());
        todo &= ~grp;
    }
    return mine;
}

__global__ void kb_cell_count(BatchDev B) {
    PARTICLE_THREAD(B);
    if (!live) return;
    uint32_t k = v.cell_base + local_key(v.grid, gm(B.ci)[3 * gi], gm(B.ci)[3 * gi + 1], gm(B.ci)[3 * gi + 2]);
    (void)wave_group_add(B.cell_cnt, k);
}

// any order within a cell: kb_cell_order sorts each cell by reference index
__global__ void kb_cell_scatter(BatchDev B) {
    PARTICLE_THREAD(B);
    if (!live) return;
    uint32_t k = v.cell_base + local_key(v.grid, gm(B.ci)[3 * gi], gm(B.ci)[3 * gi + 1], gm(B.ci)[3 * gi + 2]);
    uint32_t slot = gm(B.cell_start)[k] + wave_group_add(B.cursor, k);  // batch position
    gm(B.perm)[slot] = (uint32_t)i;                                      // scene-local source
}

// members of each cell ordered by reference index (the reference's bucket order)
__global__ void kb_cell_order(BatchDev B) {
    const unsigned lb = blockIdx.x;
    const int sid = scene_of(B.cblk, B.ns, lb);
    const SceneView &v = B.views[sid];
    const size_t c = (size_t)(lb - gm(B.cblk)[sid]) * kBlock + threadIdx.x;
    const size_t ncell = (size_t)v.grid.dim[0] * v.grid.dim[1] * v.grid.dim[2];
    if (c >= ncell) return;
    const size_t gc = v.cell_base + c;
    uint32_t lo = gm(B.cell_start)[gc], hi = gm(B.cell_start)[gc + 1];
    uint32_t *perm = B.perm;
    for (uint32_t a = lo + 1; a < hi; a++) {
        uint32_t val = perm[a];
        uint32_t key = gm(v.ref)[val];
        uint32_t b = a;
        while (b > lo) {
            uint32_t u = perm[b - 1];
            if (gm(v.ref)[u] <= key) break;
            perm[b] = u;
            b--;
        }
        perm[b] = val;
    }
}

// gather every persistent particle field into sorted order
__global__ void kb_permute(BatchDev B) {
    PARTICLE_THREAD(B);
    if (!live) return;
    const uint32_t src = gm(B.perm)[gi];
    const double x = gm(v.x)[src], y = gm(v.y)[src], z = gm(v.z)[src];
    gm(v.tx)[i] = x;
    gm(v.ty)[i] = y;
    gm(v.tz)[i] = z;
    st4(v.q4, i, x, y, z, 0.0);
    const double ux = gm(v.vx)[src], uy = gm(v.vy)[src], uz = gm(v.vz)[src];
    gm(v.tvx)[i] = ux;
    gm(v.tvy)[i] = uy;
    gm(v.tvz)[i] = uz;
    st3(v.v3, i, ux, uy, uz);
    gm(v.tm)[i] = gm(v.m)[src];
    gm(v.tfid)[i] = gm(v.fid)[src];
    gm(v.tref)[i] = gm(v.ref)[src];
    if (v.integrator == SPH_INTEGRATOR_VERLET) {  // the slot persists across steps
        gm(v.sxp)[i] = gm(v.ix)[src];
        gm(v.syp)[i] = gm(v.iy)[src];
        gm(v.szp)[i] = gm(v.iz)[src];
    }
}

// ------------------------------------------------------------------------
// neighbor search into the pair cache (spatial_hashing.rs:75-103): count,
// then indices; the 27-cell stencil is walked in the reference's (dx, dy, dz)
// order as 9 contiguous row ranges. d2 < dthr is exactly sqrt(d2) < h.
// ------------------------------------------------------------------------
__global__ void kb_count_ell(BatchDev B) {
    PARTICLE_THREAD(B);
    uint32_t cnt = 0;
    if (live) {
        const rec4 qi = ld4(v.q4, i);
        const double px = qi.x, py = qi.y, pz = qi.z;
        const GridDev &g = v.grid;
        const int cx = sat_i32(floor(px / v.cs)), cy = sat_i32(floor(py / v.cs)),
                  cz = sat_i32(floor(pz / v.cs));
        int zlo = cz - v.reach - g.org[2], zhi = cz + v.reach - g.org[2];
        if (zlo < 0) zlo = 0;
        if (zhi > g.dim[2] - 1) zhi = g.dim[2] - 1;
        if (g.part) {  // shared hash buckets: the reference's bucket visits
            bucket_walk(g, cx, cy, cz, v.reach, [&](uint32_t jg) {
                const rec4 qj = ld4(v.q4, jg - v.part_base);
                double ddx = px - qj.x, ddy = py - qj.y, ddz = pz - qj.z;
                double d2 = (ddx * ddx + ddy * ddy) + ddz * ddz;
                cnt += d2 < v.dthr ? 1u : 0u;
            });
        } else if (zlo <= zhi) {
            for (int dx = -v.reach; dx <= v.reach; dx++) {
                int gx = cx + dx - g.org[0];
                if (gx < 0 || gx >= g.dim[0]) continue;
                for (int dy = -v.reach; dy <= v.reach; dy++) {
                    int gy = cy + dy - g.org[1];
                    if (gy < 0 || gy >= g.dim[1]) continue;
                    uint32_t row = ((uint32_t)gx * (uint32_t)g.dim[1] + (uint32_t)gy) * (uint32_t)g.dim[2];
                    uint32_t a = gm(g.start)[row + zlo] - v.part_base,
                             b = gm(g.start)[row + zhi + 1] - v.part_base;
                    for (uint32_t j = a; j < b; j++) {
                        const rec4 qj = ld4(v.q4, j);
                        double ddx = px - qj.x, ddy = py - qj.y, ddz = pz - qj.z;
                        double d2 = (ddx * ddx + ddy * ddy) + ddz * ddz;
                        cnt += d2 < v.dthr ? 1u : 0u;
                    }
                }
            }
        }
        gm(B.cnt)[gi] = cnt;
    }
    uint32_t mx = cnt;
    for (int off = 32; off > 0; off >>= 1) mx = max(mx, (uint32_t)__shfl_xor((int)mx, off));
    if ((threadIdx.x & 63) == 0 && live) gm(B.slice_len)[v.slice_base + (i >> 6)] = (mx + 1) & ~1u;
}

__global__ void kb_fill_ell(BatchDev B) {
    PARTICLE_THREAD(B);
    if (!live) return;
    const rec4 qi = ld4(v.q4, i);
    const double px = qi.x, py = qi.y, pz = qi.z;
    const GridDev &g = v.grid;
    const int cx = sat_i32(floor(px / v.cs)), cy = sat_i32(floor(py / v.cs)),
              cz = sat_i32(floor(pz / v.cs));
    const size_t e0 = ELL_BASE(B, v, i);
    uint32_t kk = 0;
    if (g.part) {  // shared hash buckets: the reference's bucket visits
        bucket_walk(g, cx, cy, cz, v.reach, [&](uint32_t jg) {
            const uint32_t j = jg - v.part_base;
            const rec4 qj = ld4(v.q4, j);
            double rx = px - qj.x, ry = py - qj.y, rz = pz - qj.z;
            double d2 = (rx * rx + ry * ry) + rz * rz;
            if (d2 < v.dthr) {
                gm(B.ell_j)[ELL_AT(e0, kk)] = j;
                kk++;
            }
        });
        return;
    }
    int zlo = cz - v.reach - g.org[2], zhi = cz + v.reach - g.org[2];
    if (zlo < 0) zlo = 0;
    if (zhi > g.dim[2] - 1) zhi = g.dim[2] - 1;
    if (zlo > zhi) return;
    for (int dx = -v.reach; dx <= v.reach; dx++) {
        int gx = cx + dx - g.org[0];
        if (gx < 0 || gx >= g.dim[0]) continue;
        for (int dy = -v.reach; dy <= v.reach; dy++) {
            int gy = cy + dy - g.org[1];
            if (gy < 0 || gy >= g.dim[1]) continue;
            uint32_t row = ((uint32_t)gx * (uint32_t)g.dim[1] + (uint32_t)gy) * (uint32_t)g.dim[2];
            uint32_t a = gm(g.start)[row + zlo] - v.part_base, b = gm(g.start)[row + zhi + 1] - v.part_base;
            for (uint32_t j = a; j < b; j++) {
                const rec4 qj = ld4(v.q4, j);
                double rx = px - qj.x, ry = py - qj.y, rz = pz - qj.z;
                double d2 = (rx * rx + ry * ry) + rz * rz;
                if (d2 < v.dthr) {
                    gm(B.ell_j)[ELL_AT(e0, kk)] = j;
                    kk++;
                }
            }
        }
    }
}

// ------------------------------------------------------------------------
// One sweep per update: neighbor search into a fixed-width pair cache plus
// get_volume (spatial_hashing.rs:75-103; quantities.rs:14-61). Slice g owns
// slots [64 cap g, 64 cap (g+1)); the k-th neighbor of lane l sits at
// 64 cap g + 64 k + l. Each wavefront collects its first kSweepRows neighbor
// rows in LDS and writes them out as whole 256-byte rows once its sweep is
// done; neighbors past kSweepRows go straight to the cache, and past `cap`
// only count (the caller widens the cache and sweeps again). The volume sum
// takes the neighbors in the order they are found, which is the order every
// other pass reads them in, with kb_pairs_volume's expressions; the boundary
// terms follow as in kb_pairs_volume, so boundary samples must be current.
// The particle itself is a neighbor of the reference's list (r = 0); its
// volume term W(0) is summed in its place, but it is not stored: in every
// pass that reads the cache its term is grad W(0) = +0 times finite factors,
// a zero that leaves the pass's sums unchanged (each sum starts at +0 and a
// round-to-nearest sum is never -0, so adding a zero of either sign is the
// identity), so one slot per particle is saved in every such pass.
// ------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void kb_sweep_ell(BatchDev B, uint32_t cap, uint32_t *maxc) {
    __shared__ uint32_t rows[kBlock / 64][kSweepRows][64];
    PARTICLE_THREAD(B);
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    uint32_t cnt = 0;
    // slot of lane l's first neighbor; for lanes past the scene's end it is
    // the slot the lane would have, inside the slice's allocation
    const size_t e0 = ((size_t)(v.slice_base + (i >> 6)) * cap) * 64 + 2 * (i & 63);
    if (live) {
        const KernelConst K = v.K;
        const double rh = refined_rcp(K.h);
        const rec4 qi = ld4(v.q4, i);
        const double px = qi.x, py = qi.y, pz = qi.z;
        const GridDev &g = v.grid;
        const int cx = sat_i32(floor(px / v.cs)), cy = sat_i32(floor(py / v.cs)),
                  cz = sat_i32(floor(pz / v.cs));
        double accu = 0.;
        // a neighbor (d2 < dthr) in the order found: stored unless it is the
        // particle itself, and its volume term summed
        auto take = [&](uint32_t j, double d2) {
            if (j != i || v.keep_self) {  // the self pair is summed, not stored
                if (cnt < kSweepRows) rows[w][cnt][l] = j;
                else if (cnt < cap) gm(B.ell_j)[ELL_AT(e0, cnt)] = j;
                cnt++;
            }
            const double q = div_rcp_sum(sqrt(d2), K.h, rh);
            accu += v.V0 * (K.wfac * kernel_w(q));
        };
        int zlo = cz - v.reach - g.org[2], zhi = cz + v.reach - g.org[2];
        if (zlo < 0) zlo = 0;
        if (zhi > g.dim[2] - 1) zhi = g.dim[2] - 1;
        if (g.part) {  // shared hash buckets: the reference's bucket visits
            bucket_walk(g, cx, cy, cz, v.reach, [&](uint32_t jg) {
                const uint32_t j = jg - v.part_base;
                const GLOBAL_AS char *rec = (const GLOBAL_AS char *)v.q4 + (j << 5);
                const rec2 xy = *(const GLOBAL_AS rec2 *)rec;  // .w is being written; unused
                const double rx = px - xy.x, ry = py - xy.y,
                             rz = pz - *(const GLOBAL_AS double *)(rec + 16);
                const double d2 = (rx * rx + ry * ry) + rz * rz;
                if (d2 < v.dthr) take(j, d2);
            });
        } else if (zlo <= zhi) {
            for (int dx = -v.reach; dx <= v.reach; dx++) {
                const int gx = cx + dx - g.org[0];
                if (gx < 0 || gx >= g.dim[0]) continue;
                for (int dy = -v.reach; dy <= v.reach; dy++) {
                    const int gy = cy + dy - g.org[1];
                    if (gy < 0 || gy >= g.dim[1]) continue;
                    const uint32_t row =
                        ((uint32_t)gx * (uint32_t)g.dim[1] + (uint32_t)gy) * (uint32_t)g.dim[2];
                    const uint32_t a = gm(g.start)[row + zlo] - v.part_base,
                                   b = gm(g.start)[row + zhi + 1] - v.part_base;
                    // candidates in groups of kSweepBatch: all of a group's
                    // positions are loaded first (indices clamped to b - 1, so
                    // every load is in bounds), then tested in index order
                    for (uint32_t j0 = a; j0 < b; j0 += kSweepBatch) {
                        double cx_[kSweepBatch], cy_[kSweepBatch], cz_[kSweepBatch];
#pragma unroll
                        for (int t = 0; t < kSweepBatch; t++) {
                            const uint32_t jj = min(j0 + (uint32_t)t, b - 1);
                            const GLOBAL_AS char *rec = (const GLOBAL_AS char *)v.q4 + (jj << 5);
                            const rec2 xy = *(const GLOBAL_AS rec2 *)rec;
                            cx_[t] = xy.x;
                            cy_[t] = xy.y;
                            cz_[t] = *(const GLOBAL_AS double *)(rec + 16);  // .w is being written; unused
                        }
#pragma unroll
                        for (int t = 0; t < kSweepBatch; t++) {
                            const uint32_t j = j0 + (uint32_t)t;
                            const double rx = px - cx_[t], ry = py - cy_[t], rz = pz - cz_[t];
                            const double d2 = (rx * rx + ry * ry) + rz * rz;
                            if (j < b && d2 < v.dthr) take(j, d2);
                        }
                    }
                }
            }
        }
        for (uint32_t bi = 0; bi < v.nb; bi++) {
            const BodyDev &b = B.bodies[v.body_base + bi];
            if (b.kind == 0) {
                for (uint32_t k = gm(b.soff)[i]; k < gm(b.soff)[i + 1]; k++) {
                    const uint32_t s = gm(b.sidx)[k];
                    accu += gm(b.svol)[s] *
                            kernel_fn(K, px - gm(b.sx)[s], py - gm(b.sy)[s], pz - gm(b.sz)[s]);
                }
            } else if (gm(b.flag)[i]) {
                accu += gm(b.bvol)[i] *
                        kernel_fn(K, px - gm(b.bx)[i], py - gm(b.by)[i], pz - gm(b.bz)[i]);
            }
        }
        const double Vi = v.V0 / accu;
        gm(v.V)[i] = Vi;
        gm(v.q4)[4 * i + 3] = Vi;
        gm(B.cnt)[gi] = cnt;
    }
    uint32_t mx = cnt;
    for (int off = 32; off > 0; off >>= 1) mx = max(mx, (uint32_t)__shfl_xor((int)mx, off));
    if (l == 0 && live) {
        gm(B.slice_len)[v.slice_base + (i >> 6)] = mx;
        atomicMax(maxc, mx);  // the batch's largest count (zeroed before the launch)
    }
    // the wavefront's buffered rows as whole rows (lanes past their own count
    // write a value no pass reads)
    const uint32_t nrows = min(min(mx, (uint32_t)kSweepRows), cap);
    if (live)
        for (uint32_t k = 0; k < nrows; k++) gm(B.ell_j)[ELL_AT(e0, k)] = rows[w][k][l];
}

// slice_off[g] = g * cap: the fixed-width layout in the form ELL_BASE reads
__global__ void kb_fixed_slices(uint32_t *slice_off, size_t n, uint32_t cap) {
    const size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g <= n) slice_off[g] = (uint32_t)(g * cap);
}

// ------------------------------------------------------------------------
// VolumeMapBoundary::find_boundary_samples for every volume-map body of the
// particle's scene (volume_map_boundary.rs:1259-1424)
// ------------------------------------------------------------------------
__global__ void kb_vm_sample(BatchDev B) {
    PARTICLE_THREAD(B);
    if (!live) return;
    const double px = gm(v.x)[i], py = gm(v.y)[i], pz = gm(v.z)[i];
    const double range = v.K.h;
    for (uint32_t bi = 0; bi < v.nb; bi++) {
        const BodyDev &b = B.bodies[v.body_base + bi];
        if (b.kind != 1) continue;
        uint8_t ok = 0;
        double ox, oy, oz, vol;
        if (vm_normal_sample(b, px, py, pz, range, v.dxr, v.vmax_ball, ox, oy, oz, vol)) {
            double vx, vy, vz;
            body_velocity(b, ox, oy, oz, vx, vy, vz);
            gm(const_cast<double *>(b.bx))[i] = ox;
            gm(const_cast<double *>(b.by))[i] = oy;
            gm(const_cast<double *>(b.bz))[i] = oz;
            gm(const_cast<double *>(b.bvx))[i] = vx;
            gm(const_cast<double *>(b.bvy))[i] = vy;
            gm(const_cast<double *>(b.bvz))[i] = vz;
            gm(const_cast<double *>(b.bvol))[i] = vol;
            ok = 1;
        }
        gm(const_cast<uint8_t *>(b.flag))[i] = ok;
    }
}

// ------------------------------------------------------------------------
// grad W per cached pair (lockstep, coalesced stores) and get_volume
// (quantities.rs:14-61)
// ------------------------------------------------------------------------
template <bool CACHED, int UNROLL, int WAVES, int PIPE>
__global__ PAIR_PASS_BOUNDS void kb_pairs_volume(BatchDev B) {
    PARTICLE_THREAD(B);
    if (!live) return;
    const KernelConst K = v.K;
    const double rh = refined_rcp(K.h);  // 1/h refined once per thread
    const rec4 qi = ld4(v.q4, i);
    const double px = qi.x, py = qi.y, pz = qi.z;
    const uint32_t c = gm(B.cnt)[gi];
    const size_t e0 = ELL_BASE(B, v, i);
    double accu = 0.;
#pragma unroll UNROLL
    for (uint32_t k = 0; k < c; k++) {
        const size_t e = ELL_AT(e0, k);
        uint32_t j = gm(B.ell_j)[e];
        const rec4 qj = ldj4(v.q4, j);  // .w is being written by this pass; unused here
        double rx = px - qj.x, ry = py - qj.y, rz = pz - qj.z;
        double rn = sqrt((rx * rx + ry * ry) + rz * rz);
        double q = div_rcp_sum(rn, K.h, rh);
        accu += v.V0 * (K.wfac * kernel_w(q));
        if (CACHED) {  // the expression of kernel_grad_rh
            const double d = kernel_dqw(q);
            const double rs = rn != 0.0 ? rn : 1.0;
            const double yr = refined_rcp(rs);
            const double wx = div_rcp_sum(K.gfac * rx, rs, yr) * d;
            const double wy = div_rcp_sum(K.gfac * ry, rs, yr) * d;
            const double wz = div_rcp_sum(K.gfac * rz, rs, yr) * d;
            gm(B.ell_gx)[e] = wx;
            gm(B.ell_gy)[e] = wy;
            gm(B.ell_gz)[e] = wz;
        }
    }
    for (uint32_t bi = 0; bi < v.nb; bi++) {
        const BodyDev &b = B.bodies[v.body_base + bi];
        if (b.kind == 0) {
            for (uint32_t k = gm(b.soff)[i]; k < gm(b.soff)[i + 1]; k++) {
                uint32_t s = gm(b.sidx)[k];
                accu += gm(b.svol)[s] * kernel_fn(K, px - gm(b.sx)[s], py - gm(b.sy)[s], pz - gm(b.sz)[s]);
            }
        } else if (gm(b.flag)[i]) {
            accu += gm(b.bvol)[i] * kernel_fn(K, px - gm(b.bx)[i], py - gm(b.by)[i], pz - gm(b.bz)[i]);
        }
    }
    const double Vi = v.V0 / accu;
    gm(v.V)[i] = Vi;
    gm(v.q4)[4 * i + 3] = Vi;
}

// q4[i].w = V[i] after V was set from outside the volume pass
__global__ void kb_pack_volume(BatchDev B) {
    PARTICLE_THREAD(B);
    if (live) gm(v.q4)[4 * i + 3] = gm(v.V)[i];
}

// ------------------------------------------------------------------------
// reset + gravity + artificial viscosity (non_pressure_accelerations.rs:19-185);
// reactions onto dynamic bodies summed per workgroup in FP64
// ------------------------------------------------------------------------
template <bool CACHED, int UNROLL, int WAVES, int PIPE>
__global__ PAIR_PASS_BOUNDS void kb_nonpressure(BatchDev B) {
    __shared__ double lds[6 * kBlock];
    PARTICLE_THREAD(B);
    const KernelConst K = v.K;
    const double rh = refined_rcp(K.h);  // 1/h refined once per thread
    double px = 0, py = 0, pz = 0, ux = 0, uy = 0, uz = 0, mi = 0;
    double gx = 0.0, gy = 0.0, gz = 0.0;
    double acx = 0., acy = 0., acz = 0.;
    if (live) {
        const rec4 qi = ld4(v.q4, i);
        px = qi.x; py = qi.y; pz = qi.z;
        ux = gm(v.vx)[i]; uy = gm(v.vy)[i]; uz = gm(v.vz)[i];
        mi = gm(v.m)[i];
        if (v.gravity_mode == 0) {
            gx = 0.0 + 0.0;
            gy = 0.0 + 0.0;
            gz = 0.0 + -9.81;
        } else {
            double dx = gm(v.gcenter)[0] - px, dy = gm(v.gcenter)[1] - py, dz = gm(v.gcenter)[2] - pz;
            double dist = sqrt(dot3(dx, dy, dz, dx, dy, dz) + 1e-4);
            gx = 0.0 + (9.81 * dx) / dist;
            gy = 0.0 + (9.81 * dy) / dist;
            gz = 0.0 + (9.81 * dz) / dist;
        }
        const uint32_t c = gm(B.cnt)[gi];
        const size_t e0 = ELL_BASE(B, v, i);
        const double *q4 = v.q4, *v3 = v.v3;  // bases read once, ahead of the loop
        neighbor_loop<PIPE, UNROLL>(
            gm(B.ell_j) + e0, c, [&](uint32_t j) { return QV{ldj4(q4, j), ldj3(v3, j)}; },
            [&](uint32_t k, const QV &r) {
                double rx = px - r.q.x, ry = py - r.q.y, rz = pz - r.q.z;
                double cf = v.nu_f * 2. * (3. + 2.) * r.q.w *
                            dot3(ux - r.u.x, uy - r.u.y, uz - r.u.z, rx, ry, rz) /
                            (dot3(rx, ry, rz, rx, ry, rz) + v.eps2);
                double wx, wy, wz;
                pair_grad<CACHED>(B, K, rh, ELL_AT(e0, k), rx, ry, rz, wx, wy, wz);
                acx += cf * wx;
                acy += cf * wy;
                acz += cf * wz;
            });
    }
    for (uint32_t bi = 0; bi < v.nb; bi++) {
        const BodyDev &b = B.bodies[v.body_base + bi];
        double f[6] = {0., 0., 0., 0., 0., 0.};
        if (live) {
            if (b.kind == 0) {
                for (uint32_t e = gm(b.soff)[i]; e < gm(b.soff)[i + 1]; e++) {
                    uint32_t s = gm(b.sidx)[e];
                    const double sxk = gm(b.sx)[s], syk = gm(b.sy)[s], szk = gm(b.sz)[s];
                    double rx = px - sxk, ry = py - syk, rz = pz - szk;
                    double wx, wy, wz;
                    kernel_grad_rh(K, rh, rx, ry, rz, wx, wy, wz);
                    // samples move only with a rigid body (zero velocity otherwise)
                    const double vbx = b.dynamic ? gm(b.svx)[s] : 0.0,
                                 vby = b.dynamic ? gm(b.svy)[s] : 0.0,
                                 vbz = b.dynamic ? gm(b.svz)[s] : 0.0;
                    double cf = v.nu_b * 2. * (3. + 2.) * gm(b.svol)[s] *
                                dot3(ux - vbx, uy - vby, uz - vbz, rx, ry, rz) /
                                (dot3(rx, ry, rz, rx, ry, rz) + v.eps2);
                    const double a0 = cf * wx, a1 = cf * wy, a2 = cf * wz;
                    if (b.dynamic) {  // the reaction onto the body at the sample
                        double fx = -(mi * a0), fy = -(mi * a1), fz = -(mi * a2);
                        double tx, ty, tz;
                        cross3(sxk - gm(b.com)[0], syk - gm(b.com)[1], szk - gm(b.com)[2], fx, fy,
                               fz, tx, ty, tz);
                        f[0] += fx; f[1] += fy; f[2] += fz;
                        f[3] += tx; f[4] += ty; f[5] += tz;
                    }
                    acx += a0;
                    acy += a1;
                    acz += a2;
                }
            } else if (gm(b.flag)[i]) {
                double vp[4][3];
                visc_samples(px, py, pz, gm(b.bx)[i], gm(b.by)[i], gm(b.bz)[i], K.h, vp);
                double vv = 0.25 * gm(b.bvol)[i];
                for (int k = 0; k < 4; k++) {
                    double sxk = vp[k][0], syk = vp[k][1], szk = vp[k][2];
                    double svx, svy, svz;
                    body_velocity(b, sxk, syk, szk, svx, svy, svz);
                    double rx = px - sxk, ry = py - syk, rz = pz - szk;
                    double wx, wy, wz;
                    kernel_grad_rh(K, rh, rx, ry, rz, wx, wy, wz);
                    double cf = v.nu_b * 2. * (3. + 2.) * vv *
                                dot3(ux - svx, uy - svy, uz - svz, rx, ry, rz) /
                                (dot3(rx, ry, rz, rx, ry, rz) + v.eps2);
                    double a0 = cf * wx, a1 = cf * wy, a2 = cf * wz;
                    if (b.dynamic) {
                        double fx = -(mi * a0), fy = -(mi * a1), fz = -(mi * a2);
                        double tx, ty, tz;
                        cross3(sxk - gm(b.com)[0], syk - gm(b.com)[1], szk - gm(b.com)[2], fx, fy, fz, tx,
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
        if (b.dynamic) {  // uniform across the workgroup (one scene per workgroup)
            block_sum<6>(f, lds);
            if (threadIdx.x == 0) {
                double *dst = B.partials + ((size_t)lb * kMaxBodiesPerScene + bi) * 6;
                for (int k = 0; k < 6; k++) dst[k] = f[k];
            }
        }
    }
    if (live) {
        gm(v.ax)[i] = gx + acx;
        gm(v.ay)[i] = gy + acy;
        gm(v.az)[i] = gz + acz;
    }
}

// SESPH state equation (sesph.rs:40-51)
__global__ void kb_sesph_pressure(BatchDev B) {
    PARTICLE_THREAD(B);
    if (live) gm(v.p)[i] = v.stiffness * fmax(v.V0 / gm(v.V)[i] - 1., 0.);
}

// SESPHwSplitting (sesph_with_splitting.rs:31-81, 97-153): the density
// predicted with v* over the fluid neighbors (the particle itself included:
// the pair cache keeps it for this solver) and the normal boundary samples,
// which mirror the particle's mass and enter with v*_i alone; then the state
// equation on V = m / rho*
__global__ void kb_sesph_split_density(BatchDev B) {
    PARTICLE_THREAD(B);
    if (!live) return;
    const KernelConst K = v.K;
    const double rh = refined_rcp(K.h);
    const double dt = v.S.dt;
    const rec4 qi = ld4(v.q4, i);
    const double px = qi.x, py = qi.y, pz = qi.z, mi = gm(v.m)[i];
    const double ux = gm(v.svx)[i], uy = gm(v.svy)[i], uz = gm(v.svz)[i];
    double accu = 0.;
    const uint32_t c = gm(B.cnt)[gi];
    const size_t e0 = ELL_BASE(B, v, i);
    for (uint32_t k = 0; k < c; k++) {
        const uint32_t j = gm(B.ell_j)[ELL_AT(e0, k)];
        const rec4 q = ldj4(v.q4, j), u = ldj4(v.sv4, j);  // {x, V}, {v*, m}
        const double rx = px - q.x, ry = py - q.y, rz = pz - q.z;
        double wx, wy, wz;
        kernel_grad_rh(K, rh, rx, ry, rz, wx, wy, wz);
        accu += u.w * kernel_fn(K, rx, ry, rz) + dt * dot3(ux - u.x, uy - u.y, uz - u.z, wx, wy, wz);
    }
    for (uint32_t bi = 0; bi < v.nb; bi++) {
        const BodyDev &b = B.bodies[v.body_base + bi];
        if (b.kind == 0) {
            for (uint32_t e = gm(b.soff)[i]; e < gm(b.soff)[i + 1]; e++) {
                const uint32_t s = gm(b.sidx)[e];
                const double rx = px - gm(b.sx)[s], ry = py - gm(b.sy)[s], rz = pz - gm(b.sz)[s];
                double wx, wy, wz;
                kernel_grad_rh(K, rh, rx, ry, rz, wx, wy, wz);
                accu += gm(b.svol)[s] * mi / v.V0 * kernel_fn(K, rx, ry, rz) +
                        dt * dot3(ux, uy, uz, wx, wy, wz);
            }
        } else if (gm(b.flag)[i]) {
            const double rx = px - gm(b.bx)[i], ry = py - gm(b.by)[i], rz = pz - gm(b.bz)[i];
            double wx, wy, wz;
            kernel_grad_rh(K, rh, rx, ry, rz, wx, wy, wz);
            accu += gm(b.bvol)[i] * mi / v.V0 * kernel_fn(K, rx, ry, rz) +
                    dt * dot3(ux, uy, uz, wx, wy, wz);
        }
    }
    const double vol = mi / accu;
    gm(v.p)[i] = v.stiffness * fmax(v.V0 / vol - 1., 0.);
}

// ------------------------------------------------------------------------
// add_pressure_acceleration (pressure_solver/mod.rs:121-195); target is the
// acceleration or the solver's scratch (IISPH pressure_acc_f). Scenes whose
// solve has converged are skipped when the target is the scratch.
// ------------------------------------------------------------------------
template <bool CACHED, int UNROLL, int WAVES, int PIPE, bool DESC>
__device__ __forceinline__ void pressure_accel_impl(const BatchDev &B, int target_scratch,
                                                    int overwrite, int mirror) {
    __shared__ double lds[6 * kBlock];
    PARTICLE_THREAD_G(B, DESC);
    if (target_scratch && !gm(B.active)[sid]) return;  // uniform per workgroup
    const KernelConst K = *(const KernelConst *)&v.K;
    const double rh = refined_rcp(K.h);  // 1/h refined once per thread
    const double gamma = v.S.gamma;
    double px = 0, py = 0, pz = 0, Vi = 0, mi = 1, pi = 0;
    double acx = 0., acy = 0., acz = 0.;
    if (live) {
        const rec4 qi = ld4(v.q4, i);
        px = qi.x; py = qi.y; pz = qi.z; Vi = qi.w;
        mi = gm(v.m)[i]; pi = gm(v.p)[i];
        const uint32_t c = gm(B.cnt)[gi];
        const size_t e0 = ELL_BASE(B, v, i);
        const double vm = Vi / mi;
        const double *q4 = v.q4, *pj = v.p;  // bases read once, ahead of the loop
        neighbor_loop<PIPE, UNROLL>(
            gm(B.ell_j) + e0, c, [&](uint32_t j) { return QP{ldj4(q4, j), ldj(pj, j)}; },
            [&](uint32_t k, const QP &r) {
                double cf = vm * r.q.w * (pi + r.p);
                double wx, wy, wz;
                pair_grad<CACHED>(B, K, rh, ELL_AT(e0, k), px - r.q.x, py - r.q.y,
                                  pz - r.q.z, wx, wy, wz);
                acx -= cf * wx;
                acy -= cf * wy;
                acz -= cf * wz;
            });
    }
    for (uint32_t bi = 0; bi < v.nb; bi++) {
        const auto &b = gm(B.bodies)[v.body_base + bi];
        double f[6] = {0., 0., 0., 0., 0., 0.};
        if (live) {
            if (b.kind == 0) {
                for (uint32_t e = gm(b.soff)[i]; e < gm(b.soff)[i + 1]; e++) {
                    uint32_t s = gm(b.sidx)[e];
                    const double sxk = gm(b.sx)[s], syk = gm(b.sy)[s], szk = gm(b.sz)[s];
                    double wx, wy, wz;
                    kernel_grad_rh(K, rh, px - sxk, py - syk, pz - szk, wx, wy, wz);
                    double cf = 2. * gamma * Vi * gm(b.svol)[s] * pi;
                    const double fx = cf * wx, fy = cf * wy, fz = cf * wz;
                    if (mirror && b.dynamic) {  // the mirrored force onto the body at the sample
                        double tx, ty, tz;
                        cross3(sxk - gm(b.com)[0], syk - gm(b.com)[1], szk - gm(b.com)[2], fx, fy,
                               fz, tx, ty, tz);
                        f[0] += fx; f[1] += fy; f[2] += fz;
                        f[3] += tx; f[4] += ty; f[5] += tz;
                    }
                    acx -= fx / mi;
                    acy -= fy / mi;
                    acz -= fz / mi;
                }
            } else if (gm(b.flag)[i]) {
                double sxk = gm(b.bx)[i], syk = gm(b.by)[i], szk = gm(b.bz)[i];
                double wx, wy, wz;
                kernel_grad_rh(K, rh, px - sxk, py - syk, pz - szk, wx, wy, wz);
                double cf = 2. * gamma * Vi * gm(b.bvol)[i] * pi;
                double fx = cf * wx, fy = cf * wy, fz = cf * wz;
                if (mirror && b.dynamic) {
                    double tx, ty, tz;
                    cross3(sxk - gm(b.com)[0], syk - gm(b.com)[1], szk - gm(b.com)[2], fx, fy, fz, tx, ty,
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
                auto dst = gm(B.partials) + ((size_t)lb * kMaxBodiesPerScene + bi) * 6;
                for (int k = 0; k < 6; k++) dst[k] = f[k];
            }
        }
    }
    if (live && target_scratch) {  // the solver's scratch is always overwritten
        st3(v.pa3, i, acx, acy, acz);
    } else if (live) {
        double *tx_ = v.ax, *ty_ = v.ay, *tz_ = v.az;
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

template <bool CACHED, int UNROLL, int WAVES, int PIPE>
__global__ PAIR_PASS_BOUNDS void kb_pressure_accel(BatchDev B, int target_scratch, int overwrite,
                                                   int mirror) {
    pressure_accel_impl<CACHED, UNROLL, WAVES, PIPE, false>(B, target_scratch, overwrite, mirror);
}

template <bool CACHED, int UNROLL, int WAVES, int PIPE>
__global__ PAIR_PASS_BOUNDS void kb_pressure_accel_d(const BatchDev *__restrict__ B, int target_scratch,
                                                     int overwrite, int mirror) {
    const BatchDev Bl = *B;  // loaded before any store: its pointers stay global, uniform
    pressure_accel_impl<CACHED, UNROLL, WAVES, PIPE, true>(Bl, target_scratch, overwrite, mirror);
}

// v* = (v or v*) + dt a   (pressure_solver/mod.rs:90-108)
__global__ void kb_pred_vel(BatchDev B, int to_pred) {
    PARTICLE_THREAD(B);
    if (!live) return;
    const double dt = v.S.dt;
    double bx = to_pred ? gm(v.svx)[i] : gm(v.vx)[i], by = to_pred ? gm(v.svy)[i] : gm(v.vy)[i],
           bz = to_pred ? gm(v.svz)[i] : gm(v.vz)[i];
    const double nx = bx + dt * gm(v.ax)[i], ny = by + dt * gm(v.ay)[i],
                 nz = bz + dt * gm(v.az)[i];
    gm(v.svx)[i] = nx;
    gm(v.svy)[i] = ny;
    gm(v.svz)[i] = nz;
    st4(v.sv4, i, nx, ny, nz, gm(v.m)[i]);  // the record the source pass gathers
}

// source term (vp or vde), diagonal element and initial pressure in one pass
// (iisph.rs:117-376, clamp_pressure = true)
template <bool CACHED, int UNROLL, int WAVES, int PIPE>
__global__ PAIR_PASS_BOUNDS void kb_source_diag(BatchDev B, int vp) {
    PARTICLE_THREAD(B);
    if (!live) return;
    const KernelConst K = v.K;
    const double rh = refined_rcp(K.h);  // 1/h refined once per thread
    const SolveConst S = v.S;
    const rec4 qi = ld4(v.q4, i);
    const double px = qi.x, py = qi.y, pz = qi.z;
    const double ux = gm(v.svx)[i], uy = gm(v.svy)[i], uz = gm(v.svz)[i];
    const double Vi = qi.w;
    const double dt = S.dt, dt2 = S.dt * S.dt;
    double src = vp ? 1. - S.V0 / Vi : 0.;
    double sfx = 0., sfy = 0., sfz = 0., sf2 = 0., sbx = 0., sby = 0., sbz = 0.;
    const uint32_t c = gm(B.cnt)[gi];
    const size_t e0 = ELL_BASE(B, v, i);
    const double *q4 = v.q4, *sv4 = v.sv4;  // bases read once, ahead of the loop
    neighbor_loop<PIPE, UNROLL>(
        gm(B.ell_j) + e0, c, [&](uint32_t j) { return QU{ldj4(q4, j), ldj4(sv4, j)}; },
        [&](uint32_t k, const QU &r) {
            double wx, wy, wz;
            pair_grad<CACHED>(B, K, rh, ELL_AT(e0, k), px - r.q.x, py - r.q.y, pz - r.q.z,
                              wx, wy, wz);
            const double Vj = r.q.w;
            src -= dt * Vj * dot3(ux - r.u.x, uy - r.u.y, uz - r.u.z, wx, wy, wz);
            sfx += Vj * wx;
            sfy += Vj * wy;
            sfz += Vj * wz;
            sf2 -= dt2 * Vi * powi2(Vj) / r.u.w * dot3(wx, wy, wz, wx, wy, wz);
        });
    for (uint32_t bi = 0; bi < v.nb; bi++) {
        const BodyDev &b = B.bodies[v.body_base + bi];
        if (b.kind == 0) {
            for (uint32_t k = gm(b.soff)[i]; k < gm(b.soff)[i + 1]; k++) {
                uint32_t s = gm(b.sidx)[k];
                double wx, wy, wz;
                kernel_grad_rh(K, rh, px - gm(b.sx)[s], py - gm(b.sy)[s], pz - gm(b.sz)[s], wx, wy, wz);
                double Vb = gm(b.svol)[s];
                // samples move only with a rigid body (zero velocity otherwise)
                const double vbx = b.dynamic ? gm(b.svx)[s] : 0.0,
                             vby = b.dynamic ? gm(b.svy)[s] : 0.0,
                             vbz = b.dynamic ? gm(b.svz)[s] : 0.0;
                src -= dt * Vb * dot3(ux - vbx, uy - vby, uz - vbz, wx, wy, wz);
                sbx += Vb * wx;
                sby += Vb * wy;
                sbz += Vb * wz;
            }
        } else if (gm(b.flag)[i]) {
            double wx, wy, wz;
            kernel_grad_rh(K, rh, px - gm(b.bx)[i], py - gm(b.by)[i], pz - gm(b.bz)[i], wx, wy, wz);
            double Vb = gm(b.bvol)[i];
            src -= dt * Vb * dot3(ux - gm(b.bvx)[i], uy - gm(b.bvy)[i], uz - gm(b.bvz)[i], wx, wy, wz);
            sbx += Vb * wx;
            sby += Vb * wy;
            sbz += Vb * wz;
        }
    }
    double ci = -Vi / gm(v.m)[i];
    double cfx = ci * (sfx + 2. * S.gamma * sbx), cfy = ci * (sfy + 2. * S.gamma * sby),
           cfz = ci * (sfz + 2. * S.gamma * sbz);
    double aii = dt2 * dot3(cfx, cfy, cfz, sfx + sbx, sfy + sby, sfz + sbz) + sf2;
    gm(v.s_f)[i] = src;
    gm(v.a_ff)[i] = aii;
    gm(v.p)[i] = (aii > S.eps || aii < -S.eps) ? fmax(S.omega * src / aii, 0.) : 0.;
}

// relaxed Jacobi update and per-workgroup predicted density error
// (iisph.rs:440-531); converged scenes are skipped
template <bool CACHED, int UNROLL, int WAVES, int PIPE, bool DESC>
__device__ __forceinline__ void jacobi_impl(const BatchDev &B) {
    __shared__ double lds[kBlock];
    PARTICLE_THREAD_G(B, DESC);
    if (!gm(B.active)[sid]) return;  // uniform per workgroup
    const KernelConst K = *(const KernelConst *)&v.K;
    const double rh = refined_rcp(K.h);  // 1/h refined once per thread
    const SolveConst S = *(const SolveConst *)&v.S;
    double err = 0.;
    if (live) {
        const rec3 ai = ld3(v.pa3, i);
        const double ux = ai.x, uy = ai.y, uz = ai.z;
        const double dt2 = S.dt * S.dt;
        double adp = 0.;
        const rec4 qi = ld4(v.q4, i);
        const double px = qi.x, py = qi.y, pz = qi.z;
        const uint32_t c = gm(B.cnt)[gi];
        const size_t e0 = ELL_BASE(B, v, i);
        const double *q4 = v.q4, *pa3 = v.pa3;  // bases read once, ahead of the loop
        neighbor_loop<PIPE, UNROLL>(
            gm(B.ell_j) + e0, c, [&](uint32_t j) { return QV{ldj4(q4, j), ldj3(pa3, j)}; },
            [&](uint32_t k, const QV &r) {
                double wx, wy, wz;
                // signed-zero form: faster here, same sums (kernel_grad_rh)
                pair_grad<CACHED, true>(B, K, rh, ELL_AT(e0, k), px - r.q.x, py - r.q.y,
                                        pz - r.q.z, wx, wy, wz);
                adp += dt2 * r.q.w * dot3(ux - r.u.x, uy - r.u.y, uz - r.u.z, wx, wy, wz);
            });
        for (uint32_t bi = 0; bi < v.nb; bi++) {
            const auto &b = gm(B.bodies)[v.body_base + bi];
            if (b.kind == 0) {
                for (uint32_t k = gm(b.soff)[i]; k < gm(b.soff)[i + 1]; k++) {
                    uint32_t s = gm(b.sidx)[k];
                    double wx, wy, wz;
                    kernel_grad_rh(K, rh, px - gm(b.sx)[s], py - gm(b.sy)[s], pz - gm(b.sz)[s], wx, wy, wz);
                    adp += dt2 * gm(b.svol)[s] * dot3(ux, uy, uz, wx, wy, wz);
                }
            } else if (gm(b.flag)[i]) {
                double wx, wy, wz;
                kernel_grad_rh(K, rh, px - gm(b.bx)[i], py - gm(b.by)[i], pz - gm(b.bz)[i], wx, wy, wz);
                adp += dt2 * gm(b.bvol)[i] * dot3(ux, uy, uz, wx, wy, wz);
            }
        }
        const double aii = gm(v.a_ff)[i], si = gm(v.s_f)[i];
        if (aii < -S.eps) {
            double pn = gm(v.p)[i] + S.omega * (si - adp) / aii;
            gm(v.p)[i] = fmax(pn, 0.);
        }
        err = si < 0. ? fabs(adp - si) : 0.;
    }
    double r[1] = {err};
    block_sum<1>(r, lds);
    if (threadIdx.x == 0) gm(B.partials)[lb] = r[0];
}

template <bool CACHED, int UNROLL, int WAVES, int PIPE>
__global__ PAIR_PASS_BOUNDS void kb_jacobi(BatchDev B) {
    jacobi_impl<CACHED, UNROLL, WAVES, PIPE, false>(B);
}

template <bool CACHED, int UNROLL, int WAVES, int PIPE>
__global__ PAIR_PASS_BOUNDS void kb_jacobi_d(const BatchDev *__restrict__ B) {
    const BatchDev Bl = *B;  // loaded before any store: its pointers stay global, uniform
    jacobi_impl<CACHED, UNROLL, WAVES, PIPE, true>(Bl);
}

// IISPHwOST: x* = x + dt v*   (iisph_optimized_source_term.rs:103-113)
__global__ void kb_pred_pos(BatchDev B) {
    PARTICLE_THREAD(B);
    if (!live) return;
    const double dt = v.S.dt;
    gm(v.sxp)[i] = gm(v.x)[i] + dt * gm(v.svx)[i];
    gm(v.syp)[i] = gm(v.y)[i] + dt * gm(v.svy)[i];
    gm(v.szp)[i] = gm(v.z)[i] + dt * gm(v.svz)[i];
}

// IISPHwOST: x* += dt^2 a, v* = v* + dt a   (iisph_optimized_source_term.rs:136-203)
__global__ void kb_ost_commit(BatchDev B) {
    PARTICLE_THREAD(B);
    if (!live) return;
    const double dt = v.S.dt, dt2 = dt * dt;
    gm(v.sxp)[i] = gm(v.sxp)[i] + dt2 * gm(v.ax)[i];
    gm(v.syp)[i] = gm(v.syp)[i] + dt2 * gm(v.ay)[i];
    gm(v.szp)[i] = gm(v.szp)[i] + dt2 * gm(v.az)[i];
    gm(v.svx)[i] = gm(v.svx)[i] + dt * gm(v.ax)[i];
    gm(v.svy)[i] = gm(v.svy)[i] + dt * gm(v.ay)[i];
    gm(v.svz)[i] = gm(v.svz)[i] + dt * gm(v.az)[i];
}

// measurement sums (sph/mod.rs:444-564) per workgroup
__global__ void kb_measure(BatchDev B) {
    __shared__ double lds[3 * kBlock];
    PARTICLE_THREAD(B);
    double r[3] = {0., 0., 0.};
    if (live) {
        double Vi = gm(v.V)[i], mi = gm(v.m)[i];
        r[0] = Vi < v.V0 ? mi / Vi : mi / v.V0;
        r[1] = Vi < v.V0 ? v.V0 / Vi : 1.;
        r[2] = 0.5 * mi * dot3(gm(v.vx)[i], gm(v.vy)[i], gm(v.vz)[i], gm(v.vx)[i], gm(v.vy)[i], gm(v.vz)[i]);
    }
    block_sum<3>(r, lds);
    if (threadIdx.x == 0)
        for (int k = 0; k < 3; k++) gm(B.partials)[(size_t)lb * 3 + k] = r[k];
}

// ------------------------------------------------------------------------
// per-scene folds (one workgroup per scene, fixed-shape trees)
// ------------------------------------------------------------------------
__global__ void kb_fold_sum(const double *partials, const uint32_t *blk, int ns, int rows,
                            double *out) {
    __shared__ double lds[kBlock];
    const int s = blockIdx.x / rows, r = blockIdx.x % rows;
    double v[1] = {strided_sum(partials + r, blk[s] + threadIdx.x, blk[s + 1], (size_t)rows)};
    block_sum<1>(v, lds);
    if (threadIdx.x == 0) out[(size_t)s * rows + r] = v[0];
}

template <bool DESC>
__device__ __forceinline__ void fold_converge_impl(const BatchDev &B, int after, int k) {
    __shared__ double lds[kBlock];
    const int sk = blockIdx.x;
    if (DESC && sk >= B.ns) return;  // grid rounded up (uniform per workgroup)
    const int s = B.smap ? (int)gm(B.smap)[sk] : sk;
    // the shape of kb_fold_sum with one row
    double r[1] = {strided_sum(B.partials, gm(B.pblk)[sk] + threadIdx.x, gm(B.pblk)[sk + 1], 1)};
    block_sum<1>(r, lds);
    if (threadIdx.x == 0) {
        const size_t slot = (size_t)k * B.ns_all + s;
        if (gm(B.active)[s]) {
            const auto &v = gm(B.views)[s];
            const double err = r[0] / (double)v.n * 100.0;
            const uint32_t it = gm(B.iters)[s] + 1;
            gm(B.iters)[s] = it;
            gm(B.chunk_err)[slot] = err;
            gm(B.chunk_ran)[slot] = 1;
            const bool done = after >= 0 ? it > (uint32_t)after : (it > 2 && err < v.target_err);
            if (done) gm(B.active)[s] = 0;
        } else {
            gm(B.chunk_ran)[slot] = 0;
        }
    }
}

__global__ void kb_fold_converge(BatchDev B, int after, int k) {
    fold_converge_impl<false>(B, after, k);
}

__global__ void kb_fold_converge_d(const BatchDev *__restrict__ B, int after, int k) {
    const BatchDev Bl = *B;
    fold_converge_impl<true>(Bl, after, k);
}

__global__ void kb_fold_speed(const double *partials, const uint32_t *blk, int ns, double *out) {
    __shared__ double lds[2][kBlock];
    const int s = blockIdx.x;
    double sp = 0.0, b = 0.0;
    for (uint32_t q = blk[s] + threadIdx.x; q < blk[s + 1]; q += kBlock) {
        sp = fmax(sp, partials[(size_t)q * 2]);
        b += partials[(size_t)q * 2 + 1];
    }
    lds[0][threadIdx.x] = sp;
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
        out[2 * s] = lds[0][0];
        out[2 * s + 1] = lds[1][0];
    }
}

__global__ void kb_fold_bbox(const int *partials, const uint32_t *blk, int ns, int *out) {
    __shared__ int lds[6][kBlock];
    const int s = blockIdx.x;
    int c[6] = {2147483647, 2147483647, 2147483647, -2147483647 - 1, -2147483647 - 1,
                -2147483647 - 1};
    for (uint32_t q = blk[s] + threadIdx.x; q < blk[s + 1]; q += kBlock) {
        for (int k = 0; k < 3; k++) c[k] = min(c[k], partials[(size_t)q * 6 + k]);
        for (int k = 3; k < 6; k++) c[k] = max(c[k], partials[(size_t)q * 6 + k]);
    }
    for (int k = 0; k < 6; k++) lds[k][threadIdx.x] = c[k];
    __syncthreads();
    for (int t = kBlock / 2; t > 0; t >>= 1) {
        if (threadIdx.x < t) {
            for (int k = 0; k < 3; k++) lds[k][threadIdx.x] = min(lds[k][threadIdx.x], lds[k][threadIdx.x + t]);
            for (int k = 3; k < 6; k++) lds[k][threadIdx.x] = max(lds[k][threadIdx.x], lds[k][threadIdx.x + t]);
        }
        __syncthreads();
    }
    if (threadIdx.x < 6) out[6 * s + threadIdx.x] = lds[threadIdx.x][0];
}

#define INSTANTIATE_PAIR_PASSES(C, U, W, P)                                              \
    template __global__ void kb_pairs_volume<C, U, W, P>(BatchDev);                     \
    template __global__ void kb_nonpressure<C, U, W, P>(BatchDev);                      \
    template __global__ void kb_pressure_accel<C, U, W, P>(BatchDev, int, int, int);    \
    template __global__ void kb_source_diag<C, U, W, P>(BatchDev, int);                 \
    template __global__ void kb_jacobi<C, U, W, P>(BatchDev);                            \
    template __global__ void kb_pressure_accel_d<C, U, W, P>(const BatchDev *, int, int, int); \
    template __global__ void kb_jacobi_d<C, U, W, P>(const BatchDev *);
#define INSTANTIATE_BOTH(U, W, P) \
    INSTANTIATE_PAIR_PASSES(true, U, W, P) INSTANTIATE_PAIR_PASSES(false, U, W, P)
SPH_PAIR_VARIANTS(INSTANTIATE_BOTH)

}  // namespace sph
