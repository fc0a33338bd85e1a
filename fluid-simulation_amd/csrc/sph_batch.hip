// Per-step passes of System::update as multi-scene launches (gfx950).
//
// Every pass evaluates, per particle, the reference's expression in the
// reference's neighbor order in binary64 without FMA contraction, so a scene's
// results do not depend on which other scenes share the launch.
#include <hip/hip_runtime.h>

#include "sph_batch.hpp"

namespace sph {

// scene of workgroup lb: the largest s with blk[s] <= lb (block-uniform)
__device__ __forceinline__ int scene_of(const uint32_t *blk, int ns, uint32_t lb) {
    int lo = 0, hi = ns;
    while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if (blk[mid] <= lb) lo = mid;
        else hi = mid;
    }
    return lo;
}

#define PARTICLE_THREAD(B)                                                 \
    const unsigned lb = xcd_block();                                       \
    const int sid = scene_of((B).pblk, (B).ns, lb);                        \
    const SceneView &v = (B).views[sid];                                   \
    const size_t i = (size_t)(lb - (B).pblk[sid]) * kBlock + threadIdx.x;  \
    const bool live = i < v.n;                                             \
    const size_t gi = (size_t)v.part_base + i;                             \
    (void)gi;                                                              \
    (void)live

// pair-cache slot of particle i's first neighbor
#define ELL_BASE(B, v, i) ((size_t)(B).slice_off[(v).slice_base + ((i) >> 6)] * 64 + ((i)&63))

// ------------------------------------------------------------------------
// integration (integration_schemes/euler_cromer.rs:12-25)
// ------------------------------------------------------------------------
__global__ void kb_integrate_ec(BatchDev B) {
    PARTICLE_THREAD(B);
    if (!live || v.integrator != SPH_INTEGRATOR_EULER_CROMER) return;
    const double dt = v.S.dt;
    double nvx = v.vx[i] + dt * v.ax[i], nvy = v.vy[i] + dt * v.ay[i],
           nvz = v.vz[i] + dt * v.az[i];
    v.vx[i] = nvx;
    v.vy[i] = nvy;
    v.vz[i] = nvz;
    v.x[i] = v.x[i] + dt * nvx;
    v.y[i] = v.y[i] + dt * nvy;
    v.z[i] = v.z[i] + dt * nvz;
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
        sp = sqrt(dot3(v.vx[i], v.vy[i], v.vz[i], v.vx[i], v.vy[i], v.vz[i]));
        b = v.z[i] < v.disable_below ? 1.0 : 0.0;
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
        B.partials[(size_t)lb * 2] = sp;
        B.partials[(size_t)lb * 2 + 1] = b;
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
        int a = sat_i32(floor(v.x[i] / v.cs)), b = sat_i32(floor(v.y[i] / v.cs)),
            d = sat_i32(floor(v.z[i] / v.cs));
        B.ci[3 * gi] = a;
        B.ci[3 * gi + 1] = b;
        B.ci[3 * gi + 2] = d;
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

__global__ void kb_cell_count(BatchDev B) {
    PARTICLE_THREAD(B);
    if (!live) return;
    uint32_t k = v.cell_base + local_key(v.grid, B.ci[3 * gi], B.ci[3 * gi + 1], B.ci[3 * gi + 2]);
    atomicAdd(&B.cell_cnt[k], 1u);
}

__global__ void kb_cell_scatter(BatchDev B) {
    PARTICLE_THREAD(B);
    if (!live) return;
    uint32_t k = v.cell_base + local_key(v.grid, B.ci[3 * gi], B.ci[3 * gi + 1], B.ci[3 * gi + 2]);
    uint32_t slot = B.cell_start[k] + atomicAdd(&B.cursor[k], 1u);  // batch position
    B.perm[slot] = (uint32_t)i;                                      // scene-local source
}

// members of each cell ordered by reference index (the reference's bucket order)
__global__ void kb_cell_order(BatchDev B) {
    const unsigned lb = blockIdx.x;
    const int sid = scene_of(B.cblk, B.ns, lb);
    const SceneView &v = B.views[sid];
    const size_t c = (size_t)(lb - B.cblk[sid]) * kBlock + threadIdx.x;
    const size_t ncell = (size_t)v.grid.dim[0] * v.grid.dim[1] * v.grid.dim[2];
    if (c >= ncell) return;
    const size_t gc = v.cell_base + c;
    uint32_t lo = B.cell_start[gc], hi = B.cell_start[gc + 1];
    uint32_t *perm = B.perm;
    for (uint32_t a = lo + 1; a < hi; a++) {
        uint32_t val = perm[a];
        uint32_t key = v.ref[val];
        uint32_t b = a;
        while (b > lo) {
            uint32_t u = perm[b - 1];
            if (v.ref[u] <= key) break;
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
    const uint32_t src = B.perm[gi];
    v.tx[i] = v.x[src];
    v.ty[i] = v.y[src];
    v.tz[i] = v.z[src];
    v.tvx[i] = v.vx[src];
    v.tvy[i] = v.vy[src];
    v.tvz[i] = v.vz[src];
    v.tm[i] = v.m[src];
    v.tfid[i] = v.fid[src];
    v.tref[i] = v.ref[src];
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
        const double px = v.x[i], py = v.y[i], pz = v.z[i];
        const GridDev &g = v.grid;
        const int cx = sat_i32(floor(px / v.cs)), cy = sat_i32(floor(py / v.cs)),
                  cz = sat_i32(floor(pz / v.cs));
        int zlo = cz - v.reach - g.org[2], zhi = cz + v.reach - g.org[2];
        if (zlo < 0) zlo = 0;
        if (zhi > g.dim[2] - 1) zhi = g.dim[2] - 1;
        if (zlo <= zhi) {
            for (int dx = -v.reach; dx <= v.reach; dx++) {
                int gx = cx + dx - g.org[0];
                if (gx < 0 || gx >= g.dim[0]) continue;
                for (int dy = -v.reach; dy <= v.reach; dy++) {
                    int gy = cy + dy - g.org[1];
                    if (gy < 0 || gy >= g.dim[1]) continue;
                    uint32_t row = ((uint32_t)gx * (uint32_t)g.dim[1] + (uint32_t)gy) * (uint32_t)g.dim[2];
                    uint32_t a = g.start[row + zlo] - v.part_base,
                             b = g.start[row + zhi + 1] - v.part_base;
                    for (uint32_t j = a; j < b; j++) {
                        double ddx = px - v.x[j], ddy = py - v.y[j], ddz = pz - v.z[j];
                        double d2 = (ddx * ddx + ddy * ddy) + ddz * ddz;
                        cnt += d2 < v.dthr ? 1u : 0u;
                    }
                }
            }
        }
        B.cnt[gi] = cnt;
    }
    uint32_t mx = cnt;
    for (int off = 32; off > 0; off >>= 1) mx = max(mx, (uint32_t)__shfl_xor((int)mx, off));
    if ((threadIdx.x & 63) == 0 && live) B.slice_len[v.slice_base + (i >> 6)] = mx;
}

__global__ void kb_fill_ell(BatchDev B) {
    PARTICLE_THREAD(B);
    if (!live) return;
    const double px = v.x[i], py = v.y[i], pz = v.z[i];
    const GridDev &g = v.grid;
    const int cx = sat_i32(floor(px / v.cs)), cy = sat_i32(floor(py / v.cs)),
              cz = sat_i32(floor(pz / v.cs));
    size_t e = ELL_BASE(B, v, i);
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
            uint32_t a = g.start[row + zlo] - v.part_base, b = g.start[row + zhi + 1] - v.part_base;
            for (uint32_t j = a; j < b; j++) {
                double rx = px - v.x[j], ry = py - v.y[j], rz = pz - v.z[j];
                double d2 = (rx * rx + ry * ry) + rz * rz;
                if (d2 < v.dthr) {
                    B.ell_j[e] = j;
                    e += 64;
                }
            }
        }
    }
}

// ------------------------------------------------------------------------
// VolumeMapBoundary::find_boundary_samples for every volume-map body of the
// particle's scene (volume_map_boundary.rs:1259-1424)
// ------------------------------------------------------------------------
__global__ void kb_vm_sample(BatchDev B) {
    PARTICLE_THREAD(B);
    if (!live) return;
    const double px = v.x[i], py = v.y[i], pz = v.z[i];
    const double range = v.K.h;
    for (uint32_t bi = 0; bi < v.nb; bi++) {
        const BodyDev &b = B.bodies[v.body_base + bi];
        if (b.kind != 1) continue;
        uint8_t ok = 0;
        double sd, gx, gy, gz, vol;
        if (!body_sd(b, px, py, pz, sd) && sd < range && !body_sdg(b, px, py, pz, gx, gy, gz)) {
            double gn = sqrt(dot3(gx, gy, gz, gx, gy, gz));
            if (gn > 0.5 && gn < 2.0 && !body_vm(b, px, py, pz, vol) && vol > 0.) {
                vol = fmin(vol, v.vmax_ball);
                double d = fmax(sd + 0.25 * v.dxr, v.dxr);
                double ox = px - (gx / gn) * d, oy = py - (gy / gn) * d, oz = pz - (gz / gn) * d;
                double vx, vy, vz;
                body_velocity(b, ox, oy, oz, vx, vy, vz);
                const_cast<double *>(b.bx)[i] = ox;
                const_cast<double *>(b.by)[i] = oy;
                const_cast<double *>(b.bz)[i] = oz;
                const_cast<double *>(b.bvx)[i] = vx;
                const_cast<double *>(b.bvy)[i] = vy;
                const_cast<double *>(b.bvz)[i] = vz;
                const_cast<double *>(b.bvol)[i] = vol;
                ok = 1;
            }
        }
        const_cast<uint8_t *>(b.flag)[i] = ok;
    }
}

// ------------------------------------------------------------------------
// grad W per cached pair (lockstep, coalesced stores) and get_volume
// (quantities.rs:14-61)
// ------------------------------------------------------------------------
__global__ void kb_pairs_volume(BatchDev B) {
    PARTICLE_THREAD(B);
    if (!live) return;
    const KernelConst K = v.K;
    const double px = v.x[i], py = v.y[i], pz = v.z[i];
    const uint32_t c = B.cnt[gi];
    size_t e = ELL_BASE(B, v, i);
    double accu = 0.;
#pragma unroll 4
    for (uint32_t k = 0; k < c; k++, e += 64) {
        uint32_t j = B.ell_j[e];
        double rx = px - v.x[j], ry = py - v.y[j], rz = pz - v.z[j];
        double rn = sqrt((rx * rx + ry * ry) + rz * rz);
        double q = rn / K.h;
        accu += v.V0 * (K.wfac * kernel_w(q));
        double wx = 0.0, wy = 0.0, wz = 0.0;
        if (rn != 0.0) {
            double d = kernel_dqw(q);
            wx = ((K.gfac * rx) / rn) * d;
            wy = ((K.gfac * ry) / rn) * d;
            wz = ((K.gfac * rz) / rn) * d;
        }
        B.ell_gx[e] = wx;
        B.ell_gy[e] = wy;
        B.ell_gz[e] = wz;
    }
    for (uint32_t bi = 0; bi < v.nb; bi++) {
        const BodyDev &b = B.bodies[v.body_base + bi];
        if (b.kind == 0) {
            for (uint32_t k = b.soff[i]; k < b.soff[i + 1]; k++) {
                uint32_t s = b.sidx[k];
                accu += b.svol[s] * kernel_fn(K, px - b.sx[s], py - b.sy[s], pz - b.sz[s]);
            }
        } else if (b.flag[i]) {
            accu += b.bvol[i] * kernel_fn(K, px - b.bx[i], py - b.by[i], pz - b.bz[i]);
        }
    }
    v.V[i] = v.V0 / accu;
}

// ------------------------------------------------------------------------
// reset + gravity + artificial viscosity (non_pressure_accelerations.rs:19-185);
// reactions onto dynamic bodies summed per workgroup in FP64
// ------------------------------------------------------------------------
__global__ void kb_nonpressure(BatchDev B) {
    __shared__ double lds[6 * kBlock];
    PARTICLE_THREAD(B);
    const KernelConst K = v.K;
    double px = 0, py = 0, pz = 0, ux = 0, uy = 0, uz = 0, mi = 0;
    double gx = 0.0, gy = 0.0, gz = 0.0;
    double acx = 0., acy = 0., acz = 0.;
    if (live) {
        px = v.x[i]; py = v.y[i]; pz = v.z[i];
        ux = v.vx[i]; uy = v.vy[i]; uz = v.vz[i];
        mi = v.m[i];
        if (v.gravity_mode == 0) {
            gx = 0.0 + 0.0;
            gy = 0.0 + 0.0;
            gz = 0.0 + -9.81;
        } else {
            double dx = v.gcenter[0] - px, dy = v.gcenter[1] - py, dz = v.gcenter[2] - pz;
            double dist = sqrt(dot3(dx, dy, dz, dx, dy, dz) + 1e-4);
            gx = 0.0 + (9.81 * dx) / dist;
            gy = 0.0 + (9.81 * dy) / dist;
            gz = 0.0 + (9.81 * dz) / dist;
        }
        const uint32_t c = B.cnt[gi];
        size_t e = ELL_BASE(B, v, i);
#pragma unroll 4
        for (uint32_t k = 0; k < c; k++, e += 64) {
            uint32_t j = B.ell_j[e];
            double rx = px - v.x[j], ry = py - v.y[j], rz = pz - v.z[j];
            double cf = v.nu_f * 2. * (3. + 2.) * v.V[j] *
                        dot3(ux - v.vx[j], uy - v.vy[j], uz - v.vz[j], rx, ry, rz) /
                        (dot3(rx, ry, rz, rx, ry, rz) + v.eps2);
            acx += cf * B.ell_gx[e];
            acy += cf * B.ell_gy[e];
            acz += cf * B.ell_gz[e];
        }
    }
    for (uint32_t bi = 0; bi < v.nb; bi++) {
        const BodyDev &b = B.bodies[v.body_base + bi];
        double f[6] = {0., 0., 0., 0., 0., 0.};
        if (live) {
            if (b.kind == 0) {
                for (uint32_t e = b.soff[i]; e < b.soff[i + 1]; e++) {
                    uint32_t s = b.sidx[e];
                    double rx = px - b.sx[s], ry = py - b.sy[s], rz = pz - b.sz[s];
                    double wx, wy, wz;
                    kernel_grad(K, rx, ry, rz, wx, wy, wz);
                    // static samples have zero velocity
                    double cf = v.nu_b * 2. * (3. + 2.) * b.svol[s] *
                                dot3(ux - 0.0, uy - 0.0, uz - 0.0, rx, ry, rz) /
                                (dot3(rx, ry, rz, rx, ry, rz) + v.eps2);
                    acx += cf * wx;
                    acy += cf * wy;
                    acz += cf * wz;
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
                    double cf = v.nu_b * 2. * (3. + 2.) * vv *
                                dot3(ux - svx, uy - svy, uz - svz, rx, ry, rz) /
                                (dot3(rx, ry, rz, rx, ry, rz) + v.eps2);
                    double a0 = cf * wx, a1 = cf * wy, a2 = cf * wz;
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
        if (b.dynamic) {  // uniform across the workgroup (one scene per workgroup)
            block_sum<6>(f, lds);
            if (threadIdx.x == 0) {
                double *dst = B.partials + ((size_t)lb * kMaxBodiesPerScene + bi) * 6;
                for (int k = 0; k < 6; k++) dst[k] = f[k];
            }
        }
    }
    if (live) {
        v.ax[i] = gx + acx;
        v.ay[i] = gy + acy;
        v.az[i] = gz + acz;
    }
}

// SESPH state equation (sesph.rs:40-51)
__global__ void kb_sesph_pressure(BatchDev B) {
    PARTICLE_THREAD(B);
    if (live) v.p[i] = v.stiffness * fmax(v.V0 / v.V[i] - 1., 0.);
}

// ------------------------------------------------------------------------
// add_pressure_acceleration (pressure_solver/mod.rs:121-195); target is the
// acceleration or the solver's scratch (IISPH pressure_acc_f). Scenes whose
// solve has converged are skipped when the target is the scratch.
// ------------------------------------------------------------------------
__global__ void kb_pressure_accel(BatchDev B, int target_scratch, int overwrite, int mirror) {
    __shared__ double lds[6 * kBlock];
    PARTICLE_THREAD(B);
    if (target_scratch && !v.active) return;  // uniform per workgroup
    const KernelConst K = v.K;
    const double gamma = v.S.gamma;
    double px = 0, py = 0, pz = 0, Vi = 0, mi = 1, pi = 0;
    double acx = 0., acy = 0., acz = 0.;
    if (live) {
        Vi = v.V[i]; mi = v.m[i]; pi = v.p[i];
        const uint32_t c = B.cnt[gi];
        size_t e = ELL_BASE(B, v, i);
        const double vm = Vi / mi;
#pragma unroll 4
        for (uint32_t k = 0; k < c; k++, e += 64) {
            uint32_t j = B.ell_j[e];
            double cf = vm * v.V[j] * (pi + v.p[j]);
            acx -= cf * B.ell_gx[e];
            acy -= cf * B.ell_gy[e];
            acz -= cf * B.ell_gz[e];
        }
        if (v.nb) {
            px = v.x[i]; py = v.y[i]; pz = v.z[i];
        }
    }
    for (uint32_t bi = 0; bi < v.nb; bi++) {
        const BodyDev &b = B.bodies[v.body_base + bi];
        double f[6] = {0., 0., 0., 0., 0., 0.};
        if (live) {
            if (b.kind == 0) {
                for (uint32_t e = b.soff[i]; e < b.soff[i + 1]; e++) {
                    uint32_t s = b.sidx[e];
                    double wx, wy, wz;
                    kernel_grad(K, px - b.sx[s], py - b.sy[s], pz - b.sz[s], wx, wy, wz);
                    double cf = 2. * gamma * Vi * b.svol[s] * pi;
                    acx -= (cf * wx) / mi;
                    acy -= (cf * wy) / mi;
                    acz -= (cf * wz) / mi;
                }
            } else if (b.flag[i]) {
                double sxk = b.bx[i], syk = b.by[i], szk = b.bz[i];
                double wx, wy, wz;
                kernel_grad(K, px - sxk, py - syk, pz - szk, wx, wy, wz);
                double cf = 2. * gamma * Vi * b.bvol[i] * pi;
                double fx = cf * wx, fy = cf * wy, fz = cf * wz;
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
                double *dst = B.partials + ((size_t)lb * kMaxBodiesPerScene + bi) * 6;
                for (int k = 0; k < 6; k++) dst[k] = f[k];
            }
        }
    }
    if (live) {
        double *tx_ = target_scratch ? v.pax : v.ax;
        double *ty_ = target_scratch ? v.pay : v.ay;
        double *tz_ = target_scratch ? v.paz : v.az;
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

// v* = (v or v*) + dt a   (pressure_solver/mod.rs:90-108)
__global__ void kb_pred_vel(BatchDev B, int to_pred) {
    PARTICLE_THREAD(B);
    if (!live) return;
    const double dt = v.S.dt;
    double bx = to_pred ? v.svx[i] : v.vx[i], by = to_pred ? v.svy[i] : v.vy[i],
           bz = to_pred ? v.svz[i] : v.vz[i];
    v.svx[i] = bx + dt * v.ax[i];
    v.svy[i] = by + dt * v.ay[i];
    v.svz[i] = bz + dt * v.az[i];
}

// source term (vp or vde), diagonal element and initial pressure in one pass
// (iisph.rs:117-376, clamp_pressure = true)
__global__ void kb_source_diag(BatchDev B, int vp) {
    PARTICLE_THREAD(B);
    if (!live) return;
    const KernelConst K = v.K;
    const SolveConst S = v.S;
    const double px = v.x[i], py = v.y[i], pz = v.z[i];
    const double ux = v.svx[i], uy = v.svy[i], uz = v.svz[i];
    const double Vi = v.V[i];
    const double dt = S.dt, dt2 = S.dt * S.dt;
    double src = vp ? 1. - S.V0 / Vi : 0.;
    double sfx = 0., sfy = 0., sfz = 0., sf2 = 0., sbx = 0., sby = 0., sbz = 0.;
    const uint32_t c = B.cnt[gi];
    size_t e = ELL_BASE(B, v, i);
#pragma unroll 4
    for (uint32_t k = 0; k < c; k++, e += 64) {
        uint32_t j = B.ell_j[e];
        double wx = B.ell_gx[e], wy = B.ell_gy[e], wz = B.ell_gz[e];
        double Vj = v.V[j];
        src -= dt * Vj * dot3(ux - v.svx[j], uy - v.svy[j], uz - v.svz[j], wx, wy, wz);
        sfx += Vj * wx;
        sfy += Vj * wy;
        sfz += Vj * wz;
        sf2 -= dt2 * Vi * powi2(Vj) / v.m[j] * dot3(wx, wy, wz, wx, wy, wz);
    }
    for (uint32_t bi = 0; bi < v.nb; bi++) {
        const BodyDev &b = B.bodies[v.body_base + bi];
        if (b.kind == 0) {
            for (uint32_t k = b.soff[i]; k < b.soff[i + 1]; k++) {
                uint32_t s = b.sidx[k];
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
    double ci = -Vi / v.m[i];
    double cfx = ci * (sfx + 2. * S.gamma * sbx), cfy = ci * (sfy + 2. * S.gamma * sby),
           cfz = ci * (sfz + 2. * S.gamma * sbz);
    double aii = dt2 * dot3(cfx, cfy, cfz, sfx + sbx, sfy + sby, sfz + sbz) + sf2;
    v.s_f[i] = src;
    v.a_ff[i] = aii;
    v.p[i] = (aii > S.eps || aii < -S.eps) ? fmax(S.omega * src / aii, 0.) : 0.;
}

// relaxed Jacobi update and per-workgroup predicted density error
// (iisph.rs:440-531); converged scenes are skipped
__global__ void kb_jacobi(BatchDev B) {
    __shared__ double lds[kBlock];
    PARTICLE_THREAD(B);
    if (!v.active) return;  // uniform per workgroup
    const KernelConst K = v.K;
    const SolveConst S = v.S;
    double err = 0.;
    if (live) {
        const double ux = v.pax[i], uy = v.pay[i], uz = v.paz[i];
        const double dt2 = S.dt * S.dt;
        double adp = 0.;
        const uint32_t c = B.cnt[gi];
        size_t e = ELL_BASE(B, v, i);
#pragma unroll 4
        for (uint32_t k = 0; k < c; k++, e += 64) {
            uint32_t j = B.ell_j[e];
            adp += dt2 * v.V[j] *
                   dot3(ux - v.pax[j], uy - v.pay[j], uz - v.paz[j], B.ell_gx[e], B.ell_gy[e],
                        B.ell_gz[e]);
        }
        double px = 0, py = 0, pz = 0;
        if (v.nb) {
            px = v.x[i]; py = v.y[i]; pz = v.z[i];
        }
        for (uint32_t bi = 0; bi < v.nb; bi++) {
            const BodyDev &b = B.bodies[v.body_base + bi];
            if (b.kind == 0) {
                for (uint32_t k = b.soff[i]; k < b.soff[i + 1]; k++) {
                    uint32_t s = b.sidx[k];
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
        const double aii = v.a_ff[i], si = v.s_f[i];
        if (aii < -S.eps) {
            double pn = v.p[i] + S.omega * (si - adp) / aii;
            v.p[i] = fmax(pn, 0.);
        }
        err = si < 0. ? fabs(adp - si) : 0.;
    }
    double r[1] = {err};
    block_sum<1>(r, lds);
    if (threadIdx.x == 0) B.partials[lb] = r[0];
}

// IISPHwOST: x* = x + dt v*   (iisph_optimized_source_term.rs:103-113)
__global__ void kb_pred_pos(BatchDev B) {
    PARTICLE_THREAD(B);
    if (!live) return;
    const double dt = v.S.dt;
    v.sxp[i] = v.x[i] + dt * v.svx[i];
    v.syp[i] = v.y[i] + dt * v.svy[i];
    v.szp[i] = v.z[i] + dt * v.svz[i];
}

// IISPHwOST: x* += dt^2 a, v* = v* + dt a   (iisph_optimized_source_term.rs:136-203)
__global__ void kb_ost_commit(BatchDev B) {
    PARTICLE_THREAD(B);
    if (!live) return;
    const double dt = v.S.dt, dt2 = dt * dt;
    v.sxp[i] = v.sxp[i] + dt2 * v.ax[i];
    v.syp[i] = v.syp[i] + dt2 * v.ay[i];
    v.szp[i] = v.szp[i] + dt2 * v.az[i];
    v.svx[i] = v.svx[i] + dt * v.ax[i];
    v.svy[i] = v.svy[i] + dt * v.ay[i];
    v.svz[i] = v.svz[i] + dt * v.az[i];
}

// measurement sums (sph/mod.rs:444-564) per workgroup
__global__ void kb_measure(BatchDev B) {
    __shared__ double lds[3 * kBlock];
    PARTICLE_THREAD(B);
    double r[3] = {0., 0., 0.};
    if (live) {
        double Vi = v.V[i], mi = v.m[i];
        r[0] = Vi < v.V0 ? mi / Vi : mi / v.V0;
        r[1] = Vi < v.V0 ? v.V0 / Vi : 1.;
        r[2] = 0.5 * mi * dot3(v.vx[i], v.vy[i], v.vz[i], v.vx[i], v.vy[i], v.vz[i]);
    }
    block_sum<3>(r, lds);
    if (threadIdx.x == 0)
        for (int k = 0; k < 3; k++) B.partials[(size_t)lb * 3 + k] = r[k];
}

// ------------------------------------------------------------------------
// per-scene folds (one workgroup per scene, fixed-shape trees)
// ------------------------------------------------------------------------
__global__ void kb_fold_sum(const double *partials, const uint32_t *blk, int ns, int rows,
                            double *out) {
    __shared__ double lds[kBlock];
    const int s = blockIdx.x / rows, r = blockIdx.x % rows;
    double acc = 0.0;
    for (uint32_t q = blk[s] + threadIdx.x; q < blk[s + 1]; q += kBlock)
        acc += partials[(size_t)q * rows + r];
    double v[1] = {acc};
    block_sum<1>(v, lds);
    if (threadIdx.x == 0) out[(size_t)s * rows + r] = v[0];
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

}  // namespace sph
