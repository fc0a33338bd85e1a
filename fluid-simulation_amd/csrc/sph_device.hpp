// Device-side math of the SPH hot path (gfx950, IEEE binary64).
//
// Every expression keeps the evaluation order of the reference's Rust code
// (left-associative products, powi(3) == (x*x)*x, nalgebra's
// dot == (ax*bx + ay*by) + az*bz) and the library is built with
// -ffp-contract=off, so each particle's sums are bit-identical to an
// un-fused binary64 evaluation in the reference's neighbor order.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sph {

constexpr int kMaxBodies = 8;
constexpr int kBlock = 256;

// Precomputed kernel constants: wfac = alpha / h^3, gfac = alpha / h^4 with
// alpha = 16 / pi (sph/kernel/cubic_spline.rs:9, kernel/mod.rs:29-45).
struct KernelConst {
    double h;
    double wfac;
    double gfac;
};

// Dense cubic-serendipity lattice on the device (utilities/discretization.rs:84-92).
// cell_ok[c] != 0 when all 32 nodes of cell c are present.
struct LatticeDev {
    double x_min[3];
    double x_max[3];
    double dx;
    int n[3];
    const double *values;
    const uint8_t *cell_ok;
};

// Per-boundary view passed to every pass (boundary_handling/mod.rs:100-119).
struct BodyDev {
    int kind;     // 0 StaticSampleBoundary, 1 VolumeMapBoundary
    int dynamic;  // RigidBodyMotion attached
    double com[3];
    double q[4];  // w, i, j, k
    double lin_vel[3];
    double ang_vel[3];
    LatticeDev sdf;
    LatticeDev vm;
    // static samples, sorted by cell; per-fluid-particle CSR into them;
    // sample velocities only for samples on a rigid body (null otherwise)
    const double *sx, *sy, *sz, *svol;
    const double *svx, *svy, *svz;
    const uint32_t *soff, *sidx;
    // volume map: per fluid particle, the normal-mode sample (if flag)
    const uint8_t *flag;
    const double *bx, *by, *bz, *bvx, *bvy, *bvz, *bvol;
};

// XCD-aware block order: the dispatcher hands workgroup b to XCD b % 8, so
// remap such that each XCD walks one contiguous range of sorted particles and
// the neighbor data its blocks gather stays in that XCD's L2.
constexpr unsigned kXcds = 8;
__device__ __forceinline__ unsigned xcd_block() {
    const unsigned nb = gridDim.x, b = blockIdx.x;
    const unsigned x = b % kXcds, k = b / kXcds;
    const unsigned q = nb / kXcds, r = nb % kXcds;
    return x * q + (x < r ? x : r) + k;
}

// The same order for a grid of at least 8 ceil(t / 8) workgroups serving t
// logical ones: workgroup b (on XCD b % 8) takes logical block xcd_block(t)
// when valid; the others hold no work. With gridDim.x == t it equals
// xcd_block(), so a launch over a rounded-up grid computes what a launch over
// exactly t workgroups does.
__device__ __forceinline__ unsigned xcd_block_of(unsigned t, bool &valid) {
    const unsigned b = blockIdx.x;
    const unsigned x = b % kXcds, k = b / kXcds;
    const unsigned q = t / kXcds, r = t % kXcds;
    valid = k < q + (x < r ? 1u : 0u);
    return x * q + (x < r ? x : r) + k;
}

__device__ __forceinline__ double powi2(double x) { return x * x; }
__device__ __forceinline__ double powi3(double x) { return (x * x) * x; }

__device__ __forceinline__ double dot3(double ax, double ay, double az, double bx, double by,
                                       double bz) {
    return (ax * bx + ay * by) + az * bz;
}

// w(q) and dw/dq (cubic_spline.rs:12-17)
__device__ __forceinline__ double kernel_w(double q) {
    return powi3(fmax(1. - q, 0.)) - 4. * powi3(fmax(0.5 - q, 0.));
}
__device__ __forceinline__ double kernel_dqw(double q) {
    return -3. * powi2(fmax(1. - q, 0.)) + 12. * powi2(fmax(0.5 - q, 0.));
}

// W(r, h) (kernel/mod.rs:29-31)
__device__ __forceinline__ double kernel_fn(const KernelConst &K, double rx, double ry,
                                            double rz) {
    double rn = sqrt(dot3(rx, ry, rz, rx, ry, rz));
    return K.wfac * kernel_w(rn / K.h);
}

// Binary64 quotients through a shared reciprocal. refined_rcp(b) and
// div_rcp(a, b, y) perform the operation sequence of the compiler's double
// division (reciprocal estimate, two Newton steps, quotient with one FMA
// residual correction; sign of a zero quotient from a * y) without its
// exponent pre-scaling and special-case fix-up, which are the identity for
// finite nonzero b with |b| and |1/b| normal and a zero or |a| > 2^-960 with
// a/b normal: the operands of the kernel gradient (|r| <= h, h > 0, and
// gfac r). For those the result is the correctly rounded a / b, so one
// refined reciprocal serves several numerators.
__device__ __forceinline__ double refined_rcp(double b) {
    double y = __builtin_amdgcn_rcp(b);
    double e = __builtin_fma(-b, y, 1.0);
    y = __builtin_fma(y, e, y);
    e = __builtin_fma(-b, y, 1.0);
    return __builtin_fma(y, e, y);
}
__device__ __forceinline__ double div_rcp(double a, double b, double y) {
    const double q = a * y;
    const double r = __builtin_fma(-b, q, a);
    const double res = __builtin_fma(r, y, q);
    return a == 0.0 ? q : res;
}

// div_rcp without the zero-quotient sign: a zero numerator gives a zero of
// either sign, every nonzero quotient is the correctly rounded one. For the
// quotients of the kernel and its gradient in the neighbor sums: |r| / h is
// >= +0 either way, and each gradient component only enters sums that start
// at +0, which no zero term changes (x + (+-0) = x for x != 0, and
// +0 + (+-0) = +0 - (+-0) = x - x = +0 in round to nearest), so the sign of a
// zero never reaches a result.
__device__ __forceinline__ double div_rcp_sum(double a, double b, double y) {
    const double q = a * y;
    const double r = __builtin_fma(-b, q, a);
    return __builtin_fma(r, y, q);
}

// grad W with 1/h refined once per thread (rh = refined_rcp(K.h)): the
// expression of kernel_grad below, bit for bit up to the sign of zero
// components (div_rcp_sum). ZS = true keeps the signs of zeros too (div_rcp):
// the same sums, and on gfx950 the form the compiler schedules the Jacobi
// pass best in (its gathers issued ahead of the arithmetic: 154 against
// 196 us per launch at 1M particles in the 4-neighbor shape).
template <bool ZS = false>
__device__ __forceinline__ void kernel_grad_rh(const KernelConst &K, double rh, double rx,
                                               double ry, double rz, double &gx, double &gy,
                                               double &gz) {
    const double rn = sqrt(dot3(rx, ry, rz, rx, ry, rz));
    const double rs = rn != 0.0 ? rn : 1.0;
    const double yr = refined_rcp(rs);
    if (ZS) {
        const double d = kernel_dqw(div_rcp(rn, K.h, rh));
        gx = div_rcp(K.gfac * rx, rs, yr) * d;
        gy = div_rcp(K.gfac * ry, rs, yr) * d;
        gz = div_rcp(K.gfac * rz, rs, yr) * d;
    } else {
        const double d = kernel_dqw(div_rcp_sum(rn, K.h, rh));
        gx = div_rcp_sum(K.gfac * rx, rs, yr) * d;
        gy = div_rcp_sum(K.gfac * ry, rs, yr) * d;
        gz = div_rcp_sum(K.gfac * rz, rs, yr) * d;
    }
}

// grad W(r, h) (kernel/mod.rs:37-45): ((gfac r) / |r|) w'(|r|/h), zero at
// r = 0. Branch-free, so that neighbor loads can be scheduled across it: at
// |r| = 0 the divisor is replaced by 1 and w'(0) = -3 + 12/4 is exactly 0, so
// the self pair (r = +0) yields +0 as the reference does; for |r| != 0 the
// expression is the reference's.
__device__ __forceinline__ void kernel_grad(const KernelConst &K, double rx, double ry,
                                            double rz, double &gx, double &gy, double &gz) {
    const double rn = sqrt(dot3(rx, ry, rz, rx, ry, rz));
    const double d = kernel_dqw(rn / K.h);
    const double rs = rn != 0.0 ? rn : 1.0;
    gx = ((K.gfac * rx) / rs) * d;
    gy = ((K.gfac * ry) / rs) * d;
    gz = ((K.gfac * rz) / rs) * d;
}

__device__ __forceinline__ void cross3(double ax, double ay, double az, double bx, double by,
                                       double bz, double &ox, double &oy, double &oz) {
    double x = ay * bz - az * by;
    double y = az * bx - ax * bz;
    double z = ax * by - ay * bx;
    ox = x;
    oy = y;
    oz = z;
}

// UnitQuaternion * Vector3 as nalgebra evaluates it: t = 2 (q_v x v);
// v' = t w + q_v x t + v
__device__ __forceinline__ void quat_rotate(double w, double qi, double qj, double qk, double vx,
                                            double vy, double vz, double &ox, double &oy,
                                            double &oz) {
    double tx, ty, tz, cx, cy, cz;
    cross3(qi, qj, qk, vx, vy, vz, tx, ty, tz);
    tx *= 2.0;
    ty *= 2.0;
    tz *= 2.0;
    cross3(qi, qj, qk, tx, ty, tz, cx, cy, cz);
    ox = tx * w + cx + vx;
    oy = ty * w + cy + vy;
    oz = tz * w + cz + vz;
}

__device__ __forceinline__ void world_to_local(const BodyDev &b, double px, double py, double pz,
                                               double &lx, double &ly, double &lz) {
    quat_rotate(b.q[0], -b.q[1], -b.q[2], -b.q[3], px - b.com[0], py - b.com[1], pz - b.com[2],
                lx, ly, lz);
}

// RigidBodyMotion::velocity_at_point (rigid_body_motion.rs:132-137); zero for static
__device__ __forceinline__ void body_velocity(const BodyDev &b, double px, double py, double pz,
                                              double &vx, double &vy, double &vz) {
    if (!b.dynamic) {
        vx = vy = vz = 0.0;
        return;
    }
    double cx, cy, cz;
    cross3(b.ang_vel[0], b.ang_vel[1], b.ang_vel[2], px - b.com[0], py - b.com[1],
           pz - b.com[2], cx, cy, cz);
    vx = b.lin_vel[0] + cx;
    vy = b.lin_vel[1] + cy;
    vz = b.lin_vel[2] + cz;
}

// Per-node lattice offsets in reference_nodes() order (discretization.rs:262-300)
__constant__ const int8_t kNodeOff[32][3] = {
    {0, 0, 0}, {3, 0, 0}, {0, 3, 0}, {3, 3, 0}, {0, 0, 3}, {3, 0, 3}, {0, 3, 3}, {3, 3, 3},
    {1, 0, 0}, {2, 0, 0}, {1, 3, 0}, {2, 3, 0}, {1, 0, 3}, {2, 0, 3}, {1, 3, 3}, {2, 3, 3},
    {0, 1, 0}, {0, 2, 0}, {3, 1, 0}, {3, 2, 0}, {0, 1, 3}, {0, 2, 3}, {3, 1, 3}, {3, 2, 3},
    {0, 0, 1}, {0, 0, 2}, {3, 0, 1}, {3, 0, 2}, {0, 3, 1}, {0, 3, 2}, {3, 3, 1}, {3, 3, 2}};

// Shape function of node i at (xi, eta, zeta) (discretization.rs:306-348).
// Reference coordinate of offset o: -1, -1/3, 1/3, 1 for o = 0..3.
__device__ __forceinline__ double ref_coord(int o) {
    return o == 0 ? -1.0 : (o == 1 ? -1. / 3. : (o == 2 ? 1. / 3. : 1.0));
}

// CubicSerendipityDiscretization::function (discretization.rs:437-483).
// Returns 0 ok, 1 OutOfBounds, 2 PrunedCell.
__device__ inline int serendipity(const LatticeDev &L, double px, double py, double pz,
                                  double &out) {
    double p[3] = {px, py, pz};
    int c[3];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        double fi = floor((p[d] - L.x_min[d]) / L.dx);
        if (p[d] < L.x_min[d] || p[d] > L.x_max[d]) return 1;
        int idx = (int)fi;
        if (idx > L.n[d] - 1) idx = L.n[d] - 1;
        c[d] = idx;
    }
    if (!L.cell_ok[((size_t)c[2] * L.n[1] + c[1]) * L.n[0] + c[0]]) return 2;
    double o0 = L.x_min[0] + (double)c[0] * L.dx;
    double o1 = L.x_min[1] + (double)c[1] * L.dx;
    double o2 = L.x_min[2] + (double)c[2] * L.dx;
    double xi = 2.0 * (px - o0) / L.dx - 1.0;
    double eta = 2.0 * (py - o1) / L.dx - 1.0;
    double zeta = 2.0 * (pz - o2) / L.dx - 1.0;
    double r2 = (xi * xi + eta * eta) + zeta * zeta;
    const size_t nx = 3 * (size_t)L.n[0] + 1, ny = 3 * (size_t)L.n[1] + 1;
    const size_t base = ((size_t)(3 * c[2]) * ny + (size_t)(3 * c[1])) * nx + (size_t)(3 * c[0]);
    double s = -0.0;
#pragma unroll
    for (int i = 0; i < 32; i++) {
        int ox = kNodeOff[i][0], oy = kNodeOff[i][1], oz = kNodeOff[i][2];
        double rx = ref_coord(ox), ry = ref_coord(oy), rz = ref_coord(oz);
        double v;
        if (i < 8) {
            v = (1.0 / 64.0) * (1.0 + xi * rx) * (1.0 + eta * ry) * (1.0 + zeta * rz) *
                (9.0 * r2 - 19.0);
        } else if (i < 16) {
            v = (9.0 / 64.0) * (1.0 - xi * xi) * (1.0 + 9.0 * xi * rx) * (1.0 + eta * ry) *
                (1.0 + zeta * rz);
        } else if (i < 24) {
            v = (9.0 / 64.0) * (1.0 - eta * eta) * (1.0 + 9.0 * eta * ry) * (1.0 + xi * rx) *
                (1.0 + zeta * rz);
        } else {
            v = (9.0 / 64.0) * (1.0 - zeta * zeta) * (1.0 + 9.0 * zeta * rz) * (1.0 + xi * rx) *
                (1.0 + eta * ry);
        }
        s += L.values[base + ((size_t)oz * ny + (size_t)oy) * nx + (size_t)ox] * v;
    }
    out = s;
    return 0;
}

// gradient by central differences with step dx/6 (discretization.rs:527-557)
__device__ inline int serendipity_grad(const LatticeDev &L, double px, double py, double pz,
                                       double &gx, double &gy, double &gz) {
    double h = L.dx / 6.0;
    double g[3];
    double p[3] = {px, py, pz};
#pragma unroll 1
    for (int a = 0; a < 3; a++) {
        double pp[3] = {p[0], p[1], p[2]}, pm[3] = {p[0], p[1], p[2]};
        pp[a] += h;
        pm[a] -= h;
        double fp, fm;
        int ep = serendipity(L, pp[0], pp[1], pp[2], fp);
        int em = serendipity(L, pm[0], pm[1], pm[2], fm);
        if (ep) return ep;
        if (em) return em;
        g[a] = (fp - fm) / (2.0 * h);
    }
    gx = g[0];
    gy = g[1];
    gz = g[2];
    return 0;
}


// ---- helpers shared by the kernel files ----
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

__device__ __forceinline__ int sat_i32(double v) {
    if (v != v) return 0;
    if (v >= 2147483647.0) return 2147483647;
    if (v <= -2147483648.0) return (int)(-2147483647 - 1);
    return (int)v;
}

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

// The normal-mode boundary sample of one position (volume_map_boundary.rs:
// 1259-1346): the point d = max(sd + dx/4, dx) inside the surface along the SDF
// gradient, with the volume-map value capped at the ball volume. False when the
// position has none (sd not below the range, an unusable gradient or volume).
__device__ inline bool vm_normal_sample(const BodyDev &b, double px, double py, double pz,
                                        double range, double dxr, double vmax, double &ox,
                                        double &oy, double &oz, double &vol) {
    double sd, gx, gy, gz;
    if (body_sd(b, px, py, pz, sd) || !(sd < range) || body_sdg(b, px, py, pz, gx, gy, gz))
        return false;
    const double gn = sqrt(dot3(gx, gy, gz, gx, gy, gz));
    if (!(gn > 0.5 && gn < 2.0) || body_vm(b, px, py, pz, vol) || !(vol > 0.)) return false;
    vol = fmin(vol, vmax);
    const double d = fmax(sd + 0.25 * dxr, dxr);
    ox = px - (gx / gn) * d;
    oy = py - (gy / gn) * d;
    oz = pz - (gz / gn) * d;
    return true;
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

}  // namespace sph
