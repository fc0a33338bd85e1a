// Host orchestration of the SPH hot path and its C ABI (include/sph_hip.h).
//
// A Scene is one system's device state (binary64 SoA particle fields, solver
// scratch, boundaries, rigid bodies). A Batch owns a HIP stream and the
// launch-time buffers (cell table, pair cache, sort scratch, reductions) and
// advances any number of scenes together: System::update
// (sph/mod.rs:674-731) is one sequence of multi-scene launches. A context
// (sph_ctx) is one scene stepped by its own batch; a world (sph_world) steps
// many contexts in shared launches. The host reads back only the scalars the
// reference's control flow depends on, once per phase for all scenes (max
// speed for the CFL step, the occupied-cell box, the Jacobi error per
// iteration, rigid-body forces).
#include <atomic>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <deque>
#include <limits>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/sph_hip.h"
#include "sph_batch.hpp"
#include "sph_device.hpp"
#include "sph_kernels.hpp"

namespace sph {

// ------------------------------------------------------------------------
// errors and launches
// ------------------------------------------------------------------------
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

#define HIPCK(expr)                                                                       \
    do {                                                                                  \
        hipError_t e_ = (expr);                                                           \
        if (e_ != hipSuccess)                                                             \
            throw ::sph::Error(SPH_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

#define LAUNCH(kernel, n, stream, ...)                                                    \
    do {                                                                                  \
        size_t n_ = (size_t)(n);                                                          \
        if (n_ > 0) {                                                                     \
            hipLaunchKernelGGL(kernel, dim3((unsigned)((n_ + kBlock - 1) / kBlock)),        \
                               dim3(kBlock), 0, stream, __VA_ARGS__);                     \
            HIPCK(hipGetLastError());                                                     \
        }                                                                                 \
    } while (0)

static unsigned nblocks(size_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

// ------------------------------------------------------------------------
// device buffer
// ------------------------------------------------------------------------
// device bytes the library holds in its buffers (every context, world and server)
static std::atomic<uint64_t> g_device_bytes{0};

template <class T>
struct DBuf {
    T *p = nullptr;
    size_t cap = 0;
    DBuf() = default;
    DBuf(const DBuf &) = delete;
    DBuf &operator=(const DBuf &) = delete;
    ~DBuf() { release(); }
    void reserve(size_t n) {
        if (n <= cap) return;
        size_t c = std::max<size_t>(n, cap + cap / 4);
        T *q = nullptr;
        HIPCK(hipMalloc((void **)&q, std::max<size_t>(c, 1) * sizeof(T)));
        release();
        p = q;
        cap = c;
        g_device_bytes += (uint64_t)std::max<size_t>(c, 1) * sizeof(T);
    }
    // pageable host source: the caller synchronizes before the source dies
    void upload(const T *h, size_t n, hipStream_t s) {
        reserve(n);
        if (n) HIPCK(hipMemcpyAsync(p, h, n * sizeof(T), hipMemcpyHostToDevice, s));
    }
    void download(T *h, size_t n, hipStream_t s) const {
        if (n) HIPCK(hipMemcpyAsync(h, p, n * sizeof(T), hipMemcpyDeviceToHost, s));
    }
    void copy_from(const DBuf &o, size_t n, hipStream_t s) {
        reserve(n);
        if (n) HIPCK(hipMemcpyAsync(p, o.p, n * sizeof(T), hipMemcpyDeviceToDevice, s));
    }
    void swap(DBuf &o) {
        std::swap(p, o.p);
        std::swap(cap, o.cap);
    }
    void release() {
        if (p) {
            (void)hipFree(p);
            g_device_bytes -= (uint64_t)std::max<size_t>(cap, 1) * sizeof(T);
        }
        p = nullptr;
        cap = 0;
    }
};

// ------------------------------------------------------------------------
// per-kernel timing with HIP events on a batch's stream
// ------------------------------------------------------------------------
struct ProfRec {
    int id;
    size_t units;
    hipEvent_t a, b;
};

template <int N> struct ProfT {
    bool on = false;
    using Rec = ProfRec;
    std::vector<Rec> pending;
    std::vector<hipEvent_t> pool;
    uint64_t launches[N] = {0};
    uint64_t units[N] = {0};
    double ms[N] = {0};

    hipEvent_t get() {
        if (!pool.empty()) {
            hipEvent_t e = pool.back();
            pool.pop_back();
            return e;
        }
        hipEvent_t e;
        HIPCK(hipEventCreate(&e));
        return e;
    }
    void begin(hipStream_t s, Rec &r) {
        r.a = get();
        r.b = get();
        HIPCK(hipEventRecord(r.a, s));
    }
    void end(hipStream_t s, Rec &r) {
        HIPCK(hipEventRecord(r.b, s));
        pending.push_back(r);
    }
    void drain() {  // the caller synchronized the stream
        for (auto &r : pending) {
            float t = 0;
            HIPCK(hipEventElapsedTime(&t, r.a, r.b));
            launches[r.id]++;
            units[r.id] += r.units;
            ms[r.id] += t;
            pool.push_back(r.a);
            pool.push_back(r.b);
        }
        pending.clear();
    }
    void reset() {
        for (int k = 0; k < N; k++) {
            launches[k] = units[k] = 0;
            ms[k] = 0;
        }
    }
    ~ProfT() {
        for (auto &r : pending) {
            (void)hipEventDestroy(r.a);
            (void)hipEventDestroy(r.b);
        }
        for (auto e : pool) (void)hipEventDestroy(e);
    }
};

using Prof = ProfT<SPH_PROF_KERNELS>;

// ------------------------------------------------------------------------
// host rigid body (boundary_handling/rigid_body_motion.rs), binary64
// ------------------------------------------------------------------------
struct Rigid {
    double mass = 0, Iinv_body[9] = {0}, com[3] = {0}, q[4] = {1, 0, 0, 0}, v[3] = {0},
           L[3] = {0}, F[3] = {0}, T[3] = {0}, Iinv_world[9] = {0}, w[3] = {0};

    static void quat_matrix(const double q[4], double R[9]) {  // to_rotation_matrix
        double w = q[0], i = q[1], j = q[2], k = q[3];
        double ww = w * w, ii = i * i, jj = j * j, kk = k * k;
        double ij = i * j * 2.0, wk = w * k * 2.0, wj = w * j * 2.0, ik = i * k * 2.0,
               jk = j * k * 2.0, wi = w * i * 2.0;
        R[0] = ww + ii - jj - kk; R[1] = ij - wk; R[2] = wj + ik;
        R[3] = wk + ij; R[4] = ww - ii + jj - kk; R[5] = jk - wi;
        R[6] = ik - wj; R[7] = wi + jk; R[8] = ww - ii - jj + kk;
    }
    static void mm(const double *A, const double *B, double *C) {
        double T[9];
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++)
                T[i * 3 + j] = (A[i * 3] * B[j] + A[i * 3 + 1] * B[3 + j]) + A[i * 3 + 2] * B[6 + j];
        memcpy(C, T, sizeof T);
    }
    static void mv(const double *M, const double *x, double *o) {
        double a = (M[0] * x[0] + M[1] * x[1]) + M[2] * x[2];
        double b = (M[3] * x[0] + M[4] * x[1]) + M[5] * x[2];
        double c = (M[6] * x[0] + M[7] * x[1]) + M[8] * x[2];
        o[0] = a; o[1] = b; o[2] = c;
    }
    static void tr(const double *A, double *T) {
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) T[j * 3 + i] = A[i * 3 + j];
    }
    void derive() {  // update_derived_variables (:71-76)
        double R[9], RT[9], X[9];
        quat_matrix(q, R);
        tr(R, RT);
        mm(R, Iinv_body, X);
        mm(X, RT, Iinv_world);
        mv(Iinv_world, L, w);
    }
    void init(const sph_rigid_init &r) {  // new (:38-62)
        mass = r.mass;
        memcpy(Iinv_body, r.inertia_inv_body, sizeof Iinv_body);
        memcpy(com, r.com, sizeof com);
        memcpy(q, r.q, sizeof q);
        memcpy(v, r.lin_vel, sizeof v);
        double R[9], RT[9], X[9], Iw[9];
        quat_matrix(q, R);
        tr(R, RT);
        mm(R, r.inertia_body, X);
        mm(X, RT, Iw);
        mv(Iw, r.ang_vel, L);
        for (int k = 0; k < 3; k++) F[k] = T[k] = 0;
        derive();
    }
    // local_to_world (rigid_body_motion.rs:104-107): q * p + c, with nalgebra's
    // quaternion-vector product t = 2 (q_v x p), q_v x t + t w + p
    void local_to_world(const double p[3], double o[3]) const {
        const double t0 = 2.0 * (q[2] * p[2] - q[3] * p[1]), t1 = 2.0 * (q[3] * p[0] - q[1] * p[2]),
                     t2 = 2.0 * (q[1] * p[1] - q[2] * p[0]);
        const double c0 = q[2] * t2 - q[3] * t1, c1 = q[3] * t0 - q[1] * t2, c2 = q[1] * t1 - q[2] * t0;
        o[0] = t0 * q[0] + c0 + p[0] + com[0];
        o[1] = t1 * q[0] + c1 + p[1] + com[1];
        o[2] = t2 * q[0] + c2 + p[2] + com[2];
    }
    // velocity_at_point (:132-137): v + omega x (p - c)
    void velocity_at(const double p[3], double o[3]) const {
        const double r0 = p[0] - com[0], r1 = p[1] - com[1], r2 = p[2] - com[2];
        o[0] = v[0] + (w[1] * r2 - w[2] * r1);
        o[1] = v[1] + (w[2] * r0 - w[0] * r2);
        o[2] = v[2] + (w[0] * r1 - w[1] * r0);
    }
    void add_force(const double f[3], const double loc[3]) {  // add_force (:143-146)
        for (int k = 0; k < 3; k++) F[k] += f[k];
        double r[3] = {loc[0] - com[0], loc[1] - com[1], loc[2] - com[2]};
        double t[3] = {r[1] * f[2] - r[2] * f[1], r[2] * f[0] - r[0] * f[2],
                       r[0] * f[1] - r[1] * f[0]};
        for (int k = 0; k < 3; k++) T[k] += t[k];
    }
    void add_force_torque(const double f[6]) {
        for (int k = 0; k < 3; k++) {
            F[k] += f[k];
            T[k] += f[3 + k];
        }
    }
    void step(double dt) {  // step_forward_in_time (:183-201)
        for (int k = 0; k < 3; k++) {
            double a = F[k] / mass;
            v[k] += a * dt;
            com[k] += v[k] * dt;
            L[k] += T[k] * dt;
        }
        mv(Iinv_world, L, w);
        // from_scaled_axis(omega dt) = exp((omega dt)/2)
        double hx = (w[0] * dt) / 2.0, hy = (w[1] * dt) / 2.0, hz = (w[2] * dt) / 2.0;
        double nn = (hx * hx + hy * hy) + hz * hz;
        const double eps = 2.220446049250313e-16;
        double d[4] = {1, 0, 0, 0};
        if (!(nn <= eps * eps)) {
            double n = std::sqrt(nn), s = 1.0 * std::sin(n) / n;
            d[0] = 1.0 * std::cos(n);
            d[1] = hx * s;
            d[2] = hy * s;
            d[3] = hz * s;
        }
        double a[4] = {q[0], q[1], q[2], q[3]};
        q[0] = d[0] * a[0] - d[1] * a[1] - d[2] * a[2] - d[3] * a[3];
        q[1] = d[0] * a[1] + d[1] * a[0] + d[2] * a[3] - d[3] * a[2];
        q[2] = d[0] * a[2] - d[1] * a[3] + d[2] * a[0] + d[3] * a[1];
        q[3] = d[0] * a[3] + d[1] * a[2] - d[2] * a[1] + d[3] * a[0];
        for (int k = 0; k < 3; k++) F[k] = T[k] = 0;
        derive();
    }
};

// ------------------------------------------------------------------------
// boundaries
// ------------------------------------------------------------------------
struct LatticeHost {
    LatticeDev d{};
    DBuf<double> values;
    DBuf<uint8_t> cell_ok;
    size_t nodes() const {
        return (3 * (size_t)d.n[0] + 1) * (3 * (size_t)d.n[1] + 1) * (3 * (size_t)d.n[2] + 1);
    }
    size_t cells() const { return (size_t)d.n[0] * d.n[1] * d.n[2]; }
    void load(const sph_lattice &L, hipStream_t s) {
        for (int k = 0; k < 3; k++) {
            d.x_min[k] = L.x_min[k];
            d.x_max[k] = L.x_max[k];
            if (L.n[k] <= 0 || L.n[k] > (1 << 20)) throw Error(SPH_ERR_ARG, "lattice n out of range");
            d.n[k] = (int)L.n[k];
        }
        d.dx = L.dx;
        const size_t nx = 3 * (size_t)d.n[0] + 1, ny = 3 * (size_t)d.n[1] + 1;
        // a cell is evaluable iff all 32 of its nodes are present
        // (discretization.rs:455-463)
        static const int off[32][3] = {
            {0, 0, 0}, {3, 0, 0}, {0, 3, 0}, {3, 3, 0}, {0, 0, 3}, {3, 0, 3}, {0, 3, 3}, {3, 3, 3},
            {1, 0, 0}, {2, 0, 0}, {1, 3, 0}, {2, 3, 0}, {1, 0, 3}, {2, 0, 3}, {1, 3, 3}, {2, 3, 3},
            {0, 1, 0}, {0, 2, 0}, {3, 1, 0}, {3, 2, 0}, {0, 1, 3}, {0, 2, 3}, {3, 1, 3}, {3, 2, 3},
            {0, 0, 1}, {0, 0, 2}, {3, 0, 1}, {3, 0, 2}, {0, 3, 1}, {0, 3, 2}, {3, 3, 1}, {3, 3, 2}};
        std::vector<uint8_t> ok(cells(), 1);
        std::vector<double> vals(nodes());
        for (size_t k = 0; k < vals.size(); k++) vals[k] = L.present[k] ? L.values[k] : 0.0;
        for (int cz = 0; cz < d.n[2]; cz++)
            for (int cy = 0; cy < d.n[1]; cy++)
                for (int cx = 0; cx < d.n[0]; cx++) {
                    uint8_t good = 1;
                    for (int i = 0; i < 32 && good; i++) {
                        size_t k = ((size_t)(3 * cz + off[i][2]) * ny + (3 * cy + off[i][1])) * nx +
                                   (3 * cx + off[i][0]);
                        good = L.present[k] ? 1 : 0;
                    }
                    ok[((size_t)cz * d.n[1] + cy) * d.n[0] + cx] = good;
                }
        values.upload(vals.data(), vals.size(), s);
        cell_ok.upload(ok.data(), ok.size(), s);
        HIPCK(hipStreamSynchronize(s));
        d.values = values.p;
        d.cell_ok = cell_ok.p;
    }
    void copy_from(const LatticeHost &o, hipStream_t s) {
        d = o.d;
        values.copy_from(o.values, o.nodes(), s);
        cell_ok.copy_from(o.cell_ok, o.cells(), s);
        d.values = values.p;
        d.cell_ok = cell_ok.p;
    }
};

struct Body {
    int kind = 1;  // 0 samples, 1 volume map
    bool dynamic = false;
    Rigid rb;
    // static samples
    size_t ns = 0;
    std::vector<double> spos;            // caller order, [ns][3]; body frame on a rigid body
    DBuf<double> sx, sy, sz, svol;       // sorted by (cell, index)
    DBuf<double> svx, svy, svz;          // sorted sample velocities (on a rigid body only)
    DBuf<double> svol_c;                 // pseudo volumes in caller order (on a rigid body only)
    bool svol_set = false;               // svol_c holds the volumes (computed or copied)
    DBuf<uint32_t> sperm;                // sorted -> caller sample index
    DBuf<uint32_t> sgrid;                // cell start
    DBuf<uint32_t> spart;                // bucket partners of the sample cells (GridDev::part)
    GridDev sg{};
    DBuf<uint32_t> soff, sidx, scnt;     // per fluid particle CSR
    // volume map
    LatticeHost sdf, vm;
    DBuf<uint8_t> flag;
    DBuf<double> bx, by, bz, bvx, bvy, bvz, bvol;
    double applied[6] = {0, 0, 0, 0, 0, 0};
    // body-frame bounding box of the boundary's mesh (LocalAABB,
    // volume_map_boundary.rs:1137-1150); without one the boundary takes no
    // part in rigid-rigid contact
    bool has_aabb = false;
    double aabb_c[3] = {0, 0, 0}, aabb_h[3] = {0, 0, 0};

    BodyDev view() const {
        BodyDev b{};
        b.kind = kind;
        b.dynamic = dynamic ? 1 : 0;
        memcpy(b.com, rb.com, sizeof b.com);
        memcpy(b.q, rb.q, sizeof b.q);
        memcpy(b.lin_vel, rb.v, sizeof b.lin_vel);
        memcpy(b.ang_vel, rb.w, sizeof b.ang_vel);
        b.sdf = sdf.d;
        b.vm = vm.d;
        b.sx = sx.p; b.sy = sy.p; b.sz = sz.p; b.svol = svol.p;
        if (dynamic && kind == 0) {
            b.svx = svx.p; b.svy = svy.p; b.svz = svz.p;
        }
        b.soff = soff.p; b.sidx = sidx.p;
        b.flag = flag.p;
        b.bx = bx.p; b.by = by.p; b.bz = bz.p;
        b.bvx = bvx.p; b.bvy = bvy.p; b.bvz = bvz.p; b.bvol = bvol.p;
        return b;
    }
};

// least double t with sqrt(t) >= range (correctly rounded host sqrt):
// sqrt(d2) < range  <=>  d2 < t
static double distance_threshold(double range) {
    if (!(range > 0.0)) return 0.0;
    double t = range * range;
    while (t > 0.0 && std::sqrt(t) >= range) t = std::nextafter(t, 0.0);
    while (std::sqrt(t) < range) t = std::nextafter(t, INFINITY);
    return t;
}

// ------------------------------------------------------------------------
// one system
// ------------------------------------------------------------------------
struct Scene {
    sph_params P{};
    double V0 = 0, dthr = 0;
    KernelConst K{};
    int reach = 1;

    size_t n = 0;
    // persistent fields in (cell, reference index) order
    DBuf<double> x, y, z, vx, vy, vz, ax, ay, az, m, V, p;
    DBuf<uint32_t> fid, ref;
    // solver and integrator slots
    DBuf<double> s_f, a_ff, svx, svy, svz, sxp, syp, szp, ix, iy, iz, ivx, ivy, ivz;
    // packed gather records: {x, y, z, V}, {pressure acceleration},
    // {predicted velocity, m}, {velocity}
    DBuf<double> q4, pa3, sv4, v3;
    // permutation targets
    DBuf<double> tx, ty, tz, tvx, tvy, tvz, tm;
    DBuf<uint32_t> tfid, tref;
    std::vector<std::unique_ptr<Body>> bodies;

    GridDev grid{};
    int box[6] = {0, 0, 0, 0, 0, 0};
    int checked_box[6] = {1, 1, 1, 0, 0, 0};
    bool hash_box_valid = false;
    bool hash_collides = false;  // the checked box holds cells that share hash buckets
    DBuf<uint32_t> hpart;        // their bucket partners on `grid` (GridDev::part)
    // pressure-solve state of the running update
    bool active = false;
    uint32_t iter = 0;
    double err = 0;
    // placement in the batch that steps it
    uint32_t part_base = 0, slice_base = 0, cell_base = 0, body_base = 0;

    bool boundaries_initialized = false;
    bool initialized = false;
    bool in_world = false;
    double dt = 0, time = 0;
    uint64_t steps = 0;
    // fixed dt: time = time_offset + (steps - offset_steps) * dt, the
    // reference's time() without the cfl_time_step feature (sph/mod.rs:107-116);
    // CFL dt: time accumulates (sph/mod.rs:134-139)
    double time_offset = 0;
    uint64_t offset_steps = 0;
    // a scene whose step failed (in a world, the others continue): the status
    // and message; its state is not valid after the failing step
    int failed = 0;
    std::string fail_msg;
    // Jacobi iterations after which a solve that has not met its termination
    // test fails (the reference has no cap, iisph.rs:387-390)
    uint32_t max_iter = 100000u;
    uint32_t last_iter = 0, eqs1_iter = 0;
    double last_err = 0, avg_density = 0;
    std::vector<double> err_hist;
    // rigid-rigid contact of a volume-map system: the Gauss-Legendre rule of
    // the contact-cell quadrature (degree 0: off), the last step's active
    // cells (kContactRec doubles each) and NNCG iterations, device scratch
    int contact_degree = 0;
    double contact_x[kContactMaxDegree] = {0}, contact_w[kContactMaxDegree] = {0};
    std::vector<double> contact_rec;
    uint32_t contact_iters = 0;
    DBuf<ContactPairDev> cpairs;
    DBuf<BodyDev> cbodies;
    DBuf<unsigned int> cnact;
    DBuf<uint32_t> cidx;
    DBuf<double> csums;

    size_t ncells() const { return (size_t)grid.dim[0] * grid.dim[1] * grid.dim[2]; }

    SceneView view(const uint32_t *cell_start) const {
        SceneView v{};
        v.n = (uint32_t)n;
        v.part_base = part_base;
        v.slice_base = slice_base;
        v.cell_base = cell_base;
        v.body_base = body_base;
        v.nb = (uint32_t)bodies.size();
        v.integrator = P.integrator;
        v.reach = reach;
        v.gravity_mode = P.gravity_mode;
        v.K = K;
        v.S.dt = dt;
        v.S.V0 = V0;
        v.S.eps = P.min_diagonal_element;
        v.S.omega = P.relaxation_factor;
        v.S.gamma = P.boundary_pressure_acceleration_weighting;
        v.V0 = V0;
        v.cs = P.kernel_support_radius;
        v.dthr = dthr;
        v.dxr = P.rest_density_grid_spacing;
        const double r = P.kernel_support_radius;
        v.vmax_ball = 4.0 / 3.0 * 3.14159265358979323846 * ((r * r) * r);  // ball_volume
        v.disable_below = P.disable_particles_below;
        v.stiffness = P.stiffness;
        v.nu_f = P.fluid_viscosity;
        v.nu_b = P.boundary_viscosity;
        v.eps2 = 0.01 * (P.rest_density_grid_spacing * P.rest_density_grid_spacing);
        v.target_err = P.target_density_error;
        for (int k = 0; k < 3; k++) v.gcenter[k] = P.gravity_center[k];
        v.grid = grid;
        v.grid.start = cell_start ? cell_start + cell_base : nullptr;
        v.grid.mkey = ref.p;  // the reference index of each sorted particle
        v.grid.mbase = part_base;
        v.x = x.p; v.y = y.p; v.z = z.p;
        v.vx = vx.p; v.vy = vy.p; v.vz = vz.p;
        v.ax = ax.p; v.ay = ay.p; v.az = az.p;
        v.m = m.p; v.V = V.p; v.p = p.p;
        v.s_f = s_f.p; v.a_ff = a_ff.p;
        v.q4 = q4.p; v.pa3 = pa3.p; v.sv4 = sv4.p; v.v3 = v3.p;
        v.svx = svx.p; v.svy = svy.p; v.svz = svz.p;
        v.sxp = sxp.p; v.syp = syp.p; v.szp = szp.p;
        v.ix = ix.p; v.iy = iy.p; v.iz = iz.p;
        v.keep_self = P.solver == SPH_SOLVER_SESPH_SPLITTING ? 1 : 0;
        v.tx = tx.p; v.ty = ty.p; v.tz = tz.p;
        v.tvx = tvx.p; v.tvy = tvy.p; v.tvz = tvz.p; v.tm = tm.p;
        v.fid = fid.p; v.ref = ref.p; v.tfid = tfid.p; v.tref = tref.p;
        return v;
    }
};

constexpr int kMaxChunk = 16;      // Jacobi iterations launched between host waits
constexpr size_t kMaxGraphs = 64;  // executable Jacobi-chunk graphs per batch

// ------------------------------------------------------------------------
// online tuning of the pair passes' code shapes
// ------------------------------------------------------------------------
// Every shape computes the same bits, so shapes are measured on live steps.
// A tuning round lasts kPairVariants * steps_per_shape steps: step k runs the
// once-per-step passes in shape k / steps_per_shape, and the Jacobi
// iterations rotate through the shapes one iteration at a time (launched one
// per host wait, so no launch finds its solve converged), which measures every
// shape of the two iterated passes on the same fluid states. After the round
// each pass keeps the shape with the least time per particle. A batch is tuned
// on its first steps and again when its particle count has changed by 2x since.
enum { PP_VOLUME, PP_NONPRESSURE, PP_SOURCE_DIAG, PP_PRESSURE_ACCEL, PP_JACOBI, kPairPasses };

struct Tuner {
    int steps_per_shape = 1;  // 0: off
    int state = 0;            // 0 untuned, 1 measuring, 2 tuned
    int shape = 0, step_in = 0;
    unsigned rotation = 0;    // the next Jacobi iteration's shape while measuring
    size_t tuned_particles = 0;
    double ms[kPairPasses][kPairVariants];
    double units[kPairPasses][kPairVariants];
    bool measuring() const { return state == 1; }
};

// ------------------------------------------------------------------------
// launch-time buffers shared by the scenes of a batch
// ------------------------------------------------------------------------
struct Batch {
    hipStream_t s = nullptr;
    double *pinned = nullptr;  // readbacks
    size_t pinned_cap = 0;
    // small uploads are staged through pinned memory that is recycled after
    // every stream synchronization
    char *arena = nullptr;
    size_t arena_cap = 0, arena_off = 0;
    std::vector<Scene *> sc;
    // pair passes read grad W from the pair cache (24 B per slot) or recompute it
    // from positions (the default: less HBM traffic and memory, same bits)
    bool grad_cache = false;
    // neighbor list and volume from one sweep into a fixed-width pair cache
    // (kb_sweep_ell) or from the count / fill / volume passes; ell_cap is the
    // fixed width the next sweep uses (0: not sized yet)
    bool sweep = true;
    uint32_t ell_cap = 0;
    // code shape of each pair pass (an index into SPH_PAIR_VARIANTS) and the
    // online tuner that picks them from timed steps
    int variant[kPairPasses] = {0, 0, 0, 0, 0};
    Tuner tuner;
    ProfT<kPairPasses * kPairVariants> tprof;  // tuning samples by pass and shape

    DBuf<int> ci, ipart;
    DBuf<uint32_t> perm, cnt, slice_len, slice_off, ell_j, cell_cnt, cell_start, cursor;
    DBuf<uint32_t> pblk, cblk, scan_a, scan_b, scan_c;
    DBuf<double> ell_gx, ell_gy, ell_gz, partials, red;
    DBuf<uint32_t> ablk, amap, max_cnt;
    // pressure-solve state in one allocation, so that a chunk's results come
    // back in one copy: chunk errors [kMaxChunk][ns] (f64), then as u32 the
    // chunk's ran flags [kMaxChunk][ns], the active flags [ns] and iteration
    // counts [ns]
    DBuf<double> solve_io;
    double *chunk_err() const { return solve_io.p; }
    uint32_t *chunk_ran() const {
        return reinterpret_cast<uint32_t *>(solve_io.p + (size_t)kMaxChunk * sc.size());
    }
    uint32_t *d_active() const { return chunk_ran() + (size_t)kMaxChunk * sc.size(); }
    uint32_t *d_iters() const { return d_active() + sc.size(); }
    uint32_t iter_hint = 3;  // iterations of the last target-terminated solve
    // Jacobi chunks replayed from HIP graphs (on by default for contexts and
    // worlds)
    bool use_graphs = true;
    struct ChunkGraph {
        std::string key;
        hipGraphExec_t exec;
    };
    std::vector<ChunkGraph> graphs;  // most recently used last
    std::vector<std::string> seen;   // keys launched once without a graph
    DBuf<SceneView> views;
    DBuf<BodyDev> bodies;
    // the batch descriptor that graph-replayed Jacobi chunks read (never
    // reallocated while graphs exist: graphs capture its address)
    DBuf<BatchDev> desc;
    // utility scratch (standalone queries and static samples)
    DBuf<int> bbox;
    DBuf<uint32_t> ucnt, ucursor;
    DBuf<unsigned int> flag;
    Prof prof;

    std::vector<uint32_t> hp_blk, hc_blk;
    size_t total_particles = 0, total_pblocks = 0, total_cells = 0, total_slices = 0;

    int device = -1;  // the GPU the stream belongs to

    void init() {
        HIPCK(hipGetDevice(&device));
        HIPCK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        ensure_pinned(64);
    }
    void ensure_pinned(size_t n) {
        if (n <= pinned_cap) return;
        sync();
        if (pinned) HIPCK(hipHostFree(pinned));
        pinned = nullptr;
        pinned_cap = 0;
        size_t c = std::max<size_t>(n, 64);
        HIPCK(hipHostMalloc((void **)&pinned, c * sizeof(double), hipHostMallocDefault));
        pinned_cap = c;
    }
    void sync() {
        HIPCK(hipStreamSynchronize(s));
        arena_off = 0;
    }
    // asynchronous host->device copy of `bytes` from a host buffer the caller may reuse at once
    void put(void *dst, const void *src, size_t bytes) {
        if (!bytes) return;
        const size_t need = (bytes + 255) & ~(size_t)255;
        if (arena_off + need > arena_cap) {
            sync();
            if (need > arena_cap) {
                if (arena) HIPCK(hipHostFree(arena));
                arena = nullptr;
                arena_cap = 0;
                size_t c = std::max<size_t>(need, (size_t)1 << 20);
                HIPCK(hipHostMalloc((void **)&arena, c, hipHostMallocDefault));
                arena_cap = c;
            }
        }
        char *h = arena + arena_off;
        arena_off += need;
        memcpy(h, src, bytes);
        HIPCK(hipMemcpyAsync(dst, h, bytes, hipMemcpyHostToDevice, s));
    }
    // free the launch-time buffers (they are reserved again on the next use)
    void release_scratch() {
        sync();
        ci.release(); ipart.release(); perm.release(); cnt.release(); slice_len.release();
        slice_off.release(); ell_j.release(); cell_cnt.release(); cell_start.release();
        cursor.release(); scan_a.release(); scan_b.release(); scan_c.release();
        ell_gx.release(); ell_gy.release(); ell_gz.release(); partials.release();
        ucnt.release(); ucursor.release();
    }
    void drop_graphs() {
        for (auto &g : graphs) (void)hipGraphExecDestroy(g.exec);
        graphs.clear();
        seen.clear();
    }
    ~Batch() {
        if (s) (void)hipStreamSynchronize(s);
        drop_graphs();
        if (pinned) (void)hipHostFree(pinned);
        if (arena) (void)hipHostFree(arena);
        if (s) (void)hipStreamDestroy(s);
    }
};

}  // namespace sph

using namespace sph;

struct sph_ctx {
    Scene scene;
    Batch own;
};

struct sph_world {
    Batch batch;
    std::vector<sph_ctx *> members;
};

// continuous batching: FIFO admission under slots and a byte budget, admitted
// systems resident in one world per pressure solver (a world runs one solver's
// sequence of launches), finished systems queued for the caller
constexpr int kServerWorlds = SPH_SOLVER_SESPH_SPLITTING + 1;
struct sph_server {
    struct Req {
        uint64_t ticket, bytes, steps, done;
        sph_ctx *ctx;
    };
    uint64_t budget = 0;
    int max_slots = 0;
    uint64_t next_ticket = 1, used = 0, world_steps = 0;
    int slots = 0;
    std::deque<Req> queued;          // submitted, in submission order
    std::vector<Req> admitted;       // admitted, context not yet attached
    std::vector<Req> resident;       // attached to the world
    std::deque<Req> finished;        // out of the world, not yet polled
    struct Failed {
        Req req;
        int status;
    };
    std::deque<Failed> failed;       // out of the world after a failed step
    sph_world *worlds[kServerWorlds] = {};  // by pressure solver, created on first use
    std::vector<std::pair<int, int>> options;  // set on every world, also on later ones
};

namespace sph {

static thread_local std::string g_last_error;

// ------------------------------------------------------------------------
// Static scratch of the allocation-free neighbor search
// (sph_neighbors_find_noalloc): every buffer beyond the caller's inputs and
// outputs lives in one device array that each GPU holds from the first load
// of this library until the process ends. Calls that overlap on one GPU
// share it and race.
// ------------------------------------------------------------------------
constexpr size_t kNaMaxPoints = (size_t)1 << 18;  // queries, and samples, per call
constexpr size_t kNaMaxCells = (size_t)1 << 20;   // cells of the samples' occupied box
constexpr size_t na_round(size_t b) { return (b + 255) & ~(size_t)255; }
// byte offsets of the scratch regions
struct NaLayout {
    size_t bbox_s, bbox_q, flag, last, part, sci, perm, ssx, ssy, ssz, ucnt, ucur, start, scan_a,
        scan_b, scan_c, qcnt, off, bytes;
};
constexpr NaLayout na_layout() {
    NaLayout L{};
    const size_t nb1 = (kNaMaxCells + 1023) / 1024, nb2 = (nb1 + 1023) / 1024;
    size_t o = 0;
    L.bbox_s = o;  // 6 int: the samples' occupied-cell box
    L.bbox_q = o + 32;  // 6 int: the queries' box
    L.flag = o + 64;  // uint32: hash-collision flag
    o += 256;
    L.part = o; o += na_round((kNaMaxPoints + kBlock - 1) / kBlock * 6 * sizeof(int));
    L.sci = o; o += na_round(3 * kNaMaxPoints * sizeof(int));
    L.perm = o; o += na_round(kNaMaxPoints * sizeof(uint32_t));
    L.ssx = o; o += na_round(kNaMaxPoints * sizeof(double));
    L.ssy = o; o += na_round(kNaMaxPoints * sizeof(double));
    L.ssz = o; o += na_round(kNaMaxPoints * sizeof(double));
    L.ucnt = o; o += na_round(kNaMaxCells * sizeof(uint32_t));
    L.ucur = o; o += na_round(kNaMaxCells * sizeof(uint32_t));
    L.start = o; o += na_round((kNaMaxCells + 1) * sizeof(uint32_t));
    L.scan_a = o; o += na_round((nb1 + 1) * sizeof(uint32_t));
    L.scan_b = o; o += na_round((nb1 + 1) * sizeof(uint32_t));
    L.scan_c = o; o += na_round((nb2 + 2) * sizeof(uint32_t));
    L.qcnt = o; o += na_round(kNaMaxPoints * sizeof(uint32_t));
    L.off = o; o += na_round((kNaMaxPoints + 1) * sizeof(uint32_t));
    L.last = 0;
    L.bytes = o;
    return L;
}
constexpr NaLayout kNa = na_layout();
static_assert(kNa.bytes <= (size_t)25 << 20, "no-alloc scratch larger than stated (25 MiB)");

static __device__ __align__(256) unsigned char scratch[kNa.bytes];

// The array's name is a host-side stand-in: its device address on each GPU
// comes from hipGetSymbolAddress, resolved once per device at set-up
// (sph_ctx_create, with that device current) and kept here.
struct ScratchSlot {
    bool resolved = false;
    hipError_t status = hipSuccess;
    unsigned char *ptr = nullptr;
};
static std::mutex g_scratch_mu;
static std::vector<ScratchSlot> g_scratch;  // by device ordinal

static void scratch_setup(int dev) {
    if (dev < 0) return;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if ((size_t)dev >= g_scratch.size()) g_scratch.resize((size_t)dev + 1);
    ScratchSlot &sl = g_scratch[(size_t)dev];
    if (sl.resolved) return;
    void *p = nullptr;
    sl.status = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    sl.ptr = sl.status == hipSuccess ? static_cast<unsigned char *>(p) : nullptr;
    sl.resolved = true;
}

static unsigned char *scratch_device(int dev) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (dev < 0 || (size_t)dev >= g_scratch.size() || !g_scratch[(size_t)dev].resolved)
        throw Error(SPH_ERR_STATE, "static scratch not set up on this context's device");
    const ScratchSlot &sl = g_scratch[(size_t)dev];
    if (sl.status != hipSuccess || !sl.ptr)
        throw Error(SPH_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") +
                                     hipGetErrorString(sl.status));
    return sl.ptr;
}

static void set_error(const char *what) { g_last_error = what ? what : ""; }

template <class F>
static int guarded(F &&f) {
    try {
        f();
        return SPH_OK;
    } catch (const Error &e) {
        set_error(e.what());
        return e.code;
    } catch (const std::bad_alloc &) {
        set_error("host allocation failed");
        return SPH_ERR_HIP;
    } catch (const std::exception &e) {
        set_error(e.what());
        return SPH_ERR_HIP;
    }
}

static void check_device() {
    int cnt = 0;
    hipError_t e = hipGetDeviceCount(&cnt);
    if (e != hipSuccess || cnt <= 0) throw Error(SPH_ERR_NO_DEVICE, "no HIP device visible");
}

// multi-scene launch of `nblk` workgroups; timed under `pid` when profiling
// (pid == SPH_PROF_KERNELS: not timed)
#define PLAUNCH(B, pid, units, kernel, nblk, ...)                                       \
    do {                                                                                \
        if ((nblk) > 0) {                                                               \
            Prof::Rec r_{pid, (size_t)(units), nullptr, nullptr};                       \
            const bool t_ = (B).prof.on && (pid) < SPH_PROF_KERNELS;                     \
            if (t_) (B).prof.begin((B).s, r_);                                          \
            hipLaunchKernelGGL(kernel, dim3((unsigned)(nblk)), dim3(kBlock), 0, (B).s,  \
                               __VA_ARGS__);                                            \
            HIPCK(hipGetLastError());                                                   \
            if (t_) (B).prof.end((B).s, r_);                                            \
        }                                                                               \
    } while (0)

// pair-pass kernels indexed [gradient cache][code shape]
typedef void (*KernB)(BatchDev);
typedef void (*KernBI)(BatchDev, int);
typedef void (*KernBIII)(BatchDev, int, int, int);
#define PV_F(U, W, P) &kb_pairs_volume<false, U, W, P>,
#define PV_T(U, W, P) &kb_pairs_volume<true, U, W, P>,
#define NP_F(U, W, P) &kb_nonpressure<false, U, W, P>,
#define NP_T(U, W, P) &kb_nonpressure<true, U, W, P>,
#define SD_F(U, W, P) &kb_source_diag<false, U, W, P>,
#define SD_T(U, W, P) &kb_source_diag<true, U, W, P>,
#define PA_F(U, W, P) &kb_pressure_accel<false, U, W, P>,
#define PA_T(U, W, P) &kb_pressure_accel<true, U, W, P>,
#define JA_F(U, W, P) &kb_jacobi<false, U, W, P>,
#define JA_T(U, W, P) &kb_jacobi<true, U, W, P>,
static const KernB k_volume_tab[2][kPairVariants] = {{SPH_PAIR_VARIANTS(PV_F)},
                                                     {SPH_PAIR_VARIANTS(PV_T)}};
static const KernB k_nonpressure_tab[2][kPairVariants] = {{SPH_PAIR_VARIANTS(NP_F)},
                                                          {SPH_PAIR_VARIANTS(NP_T)}};
static const KernBI k_source_diag_tab[2][kPairVariants] = {{SPH_PAIR_VARIANTS(SD_F)},
                                                           {SPH_PAIR_VARIANTS(SD_T)}};
static const KernBIII k_pressure_accel_tab[2][kPairVariants] = {{SPH_PAIR_VARIANTS(PA_F)},
                                                                {SPH_PAIR_VARIANTS(PA_T)}};
static const KernB k_jacobi_tab[2][kPairVariants] = {{SPH_PAIR_VARIANTS(JA_F)},
                                                     {SPH_PAIR_VARIANTS(JA_T)}};

// launch of pair pass `pass` over all particles of the batch in the batch's
// gradient mode and the pass's code shape; timed under `pid` when profiling
// and under `pass` while tuning
template <class Tab, class... A>
static void pair_pass_grid(Batch &B, size_t grid, int pass, int pid, const Tab &tab, A... args) {
    if (grid == 0) return;
    auto kern = tab[B.grad_cache ? 1 : 0][B.variant[pass]];
    Prof::Rec r{pid, B.total_particles, nullptr, nullptr};
    // tuning cost per particle slot of this launch (a chunk over the scenes
    // still iterating covers fewer than the batch's particles)
    Prof::Rec t{pass * kPairVariants + B.variant[pass], grid * kBlock, nullptr, nullptr};
    const bool timed = B.prof.on && pid < SPH_PROF_KERNELS, tuning = B.tuner.measuring();
    if (timed) B.prof.begin(B.s, r);
    if (tuning) B.tprof.begin(B.s, t);
    hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(kBlock), 0, B.s, args...);
    HIPCK(hipGetLastError());
    if (tuning) B.tprof.end(B.s, t);
    if (timed) B.prof.end(B.s, r);
}

// pair pass over every particle of the batch
template <class Tab, class... A>
static void pair_pass(Batch &B, int pass, int pid, const Tab &tab, A... args) {
    pair_pass_grid(B, B.total_pblocks, pass, pid, tab, args...);
}

static void tune_begin_step(Batch &B, size_t particles) {
    Tuner &T = B.tuner;
    if (T.steps_per_shape <= 0 || particles == 0) return;
    const bool changed = T.state == 2 && (particles > 2 * T.tuned_particles ||
                                          2 * particles < T.tuned_particles);
    if (T.state == 0 || changed) {
        T.state = 1;
        T.shape = 0;
        T.step_in = 0;
        for (int p = 0; p < kPairPasses; p++)
            for (int v = 0; v < kPairVariants; v++) T.ms[p][v] = T.units[p][v] = 0.0;
        B.tprof.reset();
    }
    if (T.state == 1)
        for (int p = 0; p < kPairPasses; p++) B.variant[p] = T.shape;
}

static void tune_end_step(Batch &B, size_t particles) {
    Tuner &T = B.tuner;
    if (T.state != 1) return;
    B.sync();
    B.tprof.drain();
    for (int p = 0; p < kPairPasses; p++)
        for (int v = 0; v < kPairVariants; v++) {
            T.ms[p][v] += B.tprof.ms[p * kPairVariants + v];
            T.units[p][v] += (double)B.tprof.units[p * kPairVariants + v];
        }
    B.tprof.reset();
    if (++T.step_in < T.steps_per_shape) return;
    T.step_in = 0;
    if (++T.shape < kPairVariants) return;
    for (int p = 0; p < kPairPasses; p++) {
        // a shape a pass never ran in (a pass the steps did not use) is not chosen
        int best = 0;
        double cost = INFINITY;
        for (int v = 0; v < kPairVariants; v++)
            if (T.units[p][v] > 0 && T.ms[p][v] / T.units[p][v] < cost) {
                cost = T.ms[p][v] / T.units[p][v];
                best = v;
            }
        B.variant[p] = best;
    }
    T.state = 2;
    T.tuned_particles = particles;
}

// Launches of an exclusive scan of n > 0 counts into out[0..n) on stream st,
// with block sums in a (>= nb1 + 1 entries), b (>= nb1 + 1) and c (>= nb2 + 2),
// nb1 = ceil(n / 1024), nb2 = ceil(nb1 / 1024); out[n] is not written.
static void scan_launch(hipStream_t st, const uint32_t *in, size_t n, uint32_t *out, uint32_t *a,
                        uint32_t *b, uint32_t *c) {
    const size_t BS = 1024;
    size_t nb1 = (n + BS - 1) / BS;
    hipLaunchKernelGGL(k_scan_block, dim3((unsigned)nb1), dim3(512), 0, st, in, n, out, a);
    HIPCK(hipGetLastError());
    if (nb1 > 1) {
        size_t nb2 = (nb1 + BS - 1) / BS;
        hipLaunchKernelGGL(k_scan_block, dim3((unsigned)nb2), dim3(512), 0, st, a, nb1, b, c);
        HIPCK(hipGetLastError());
        if (nb2 > 1) {
            if (nb2 > BS) throw Error(SPH_ERR_UNSUPPORTED, "scan longer than 2^30 elements");
            // third level in a single workgroup; its block sum goes past the scanned entries
            hipLaunchKernelGGL(k_scan_block, dim3(1), dim3(512), 0, st, c, nb2, c, c + nb2 + 1);
            HIPCK(hipGetLastError());
            hipLaunchKernelGGL(k_scan_add, dim3((unsigned)nb2), dim3(512), 0, st, b, nb1, c);
            HIPCK(hipGetLastError());
        }
        hipLaunchKernelGGL(k_scan_add, dim3((unsigned)nb1), dim3(512), 0, st, out, n, b);
        HIPCK(hipGetLastError());
    }
}

// exclusive scan of n counts into out[0..n]; out[n] = the total. A caller that
// knows the total passes it (known_total >= 0): no read-back, no host wait.
static uint32_t scan_u32(Batch &B, const uint32_t *in, size_t n, uint32_t *out,
                         int64_t known_total = -1) {
    const size_t BS = 1024;
    if (n == 0) {
        uint32_t z = 0;
        B.put(out, &z, sizeof z);
        return 0;
    }
    size_t nb1 = (n + BS - 1) / BS, nb2 = (nb1 + BS - 1) / BS;
    B.scan_a.reserve(nb1 + 1);
    B.scan_b.reserve(nb1 + 1);
    B.scan_c.reserve(nb2 + 2);
    scan_launch(B.s, in, n, out, B.scan_a.p, B.scan_b.p, B.scan_c.p);
    if (known_total >= 0) {
        const uint32_t t32 = (uint32_t)known_total;
        B.put(out + n, &t32, sizeof t32);
        return t32;
    }
    uint32_t *last = reinterpret_cast<uint32_t *>(B.pinned);
    HIPCK(hipMemcpyAsync(&last[0], out + n - 1, sizeof(uint32_t), hipMemcpyDeviceToHost, B.s));
    HIPCK(hipMemcpyAsync(&last[1], in + n - 1, sizeof(uint32_t), hipMemcpyDeviceToHost, B.s));
    B.sync();
    uint64_t total = (uint64_t)last[0] + last[1];
    if (total > 0xffffffffull) throw Error(SPH_ERR_UNSUPPORTED, "more than 2^32 entries");
    uint32_t t32 = (uint32_t)total;
    B.put(out + n, &t32, sizeof t32);
    return t32;
}

// grid over the cell box [lo, hi] (inclusive)
static GridDev make_grid(const int lo[3], const int hi[3]) {
    GridDev g{};
    size_t cells = 1;
    for (int k = 0; k < 3; k++) {
        g.org[k] = lo[k];
        long long d = (long long)hi[k] - lo[k] + 1;
        if (d <= 0 || d > (1 << 24)) throw Error(SPH_ERR_UNSUPPORTED, "cell grid extent out of range");
        g.dim[k] = (int)d;
        cells *= (size_t)d;
    }
    if (cells > (size_t)1 << 31)
        throw Error(SPH_ERR_UNSUPPORTED, "occupied cell box too large (particles too far apart)");
    return g;
}

// Whether the cell box [lo, hi] widened by the reach holds two cells that can
// share a stencil and share a hash bucket. The answer is kept for a checked
// box (with a margin) so that a system re-checks only when it leaves it.
static bool hash_guard(Batch &B, int *checked, bool &valid, bool &collides, const int lo[3],
                       const int hi[3], int reach) {
    int elo[3], ehi[3];  // the stencil cells of the box's particles
    for (int k = 0; k < 3; k++) {
        elo[k] = lo[k] - reach;
        ehi[k] = hi[k] + reach;
    }
    if (valid) {
        bool inside = true;
        for (int k = 0; k < 3; k++)
            inside = inside && elo[k] >= checked[k] && ehi[k] <= checked[3 + k];
        if (inside) return collides;
    }
    for (int k = 0; k < 3; k++) {  // margin so that small drifts do not re-check
        elo[k] -= 4;
        ehi[k] += 4;
    }
    GridDev g = make_grid(elo, ehi);
    B.flag.reserve(4);
    LAUNCH(k_fill_u32, 4, B.s, (uint32_t *)B.flag.p, (size_t)4, 0u);
    size_t nc = (size_t)g.dim[0] * g.dim[1] * g.dim[2];
    LAUNCH(k_hash_check, nc, B.s, g, 2 * reach, B.flag.p);
    unsigned int *f = reinterpret_cast<unsigned int *>(B.pinned);
    HIPCK(hipMemcpyAsync(f, B.flag.p, sizeof *f, hipMemcpyDeviceToHost, B.s));
    B.sync();
    collides = *f != 0;
    for (int k = 0; k < 3; k++) {
        checked[k] = elo[k];
        checked[3 + k] = ehi[k];
    }
    valid = true;
    return collides;
}

// The bucket partners of g's cells (GridDev::part) into `part`; g.part stays
// null when no two cells of g that can share a stencil share a bucket.
static void hash_partners(Batch &B, GridDev &g, DBuf<uint32_t> &part, int reach) {
    g.part = nullptr;
    g.pm = 2 * reach;  // a stencil cell sharing a bucket with a cell of the box lies this near it
    const size_t nc = ((size_t)g.dim[0] + 2 * g.pm) * ((size_t)g.dim[1] + 2 * g.pm) *
                      ((size_t)g.dim[2] + 2 * g.pm);
    part.reserve(2 * nc);
    B.flag.reserve(4);
    LAUNCH(k_fill_u32, 4, B.s, (uint32_t *)B.flag.p, (size_t)4, 0u);
    LAUNCH(k_hash_partners, nc, B.s, g, 2 * reach, part.p, B.flag.p);
    unsigned int *f = reinterpret_cast<unsigned int *>(B.pinned);
    HIPCK(hipMemcpyAsync(f, B.flag.p, 2 * sizeof *f, hipMemcpyDeviceToHost, B.s));
    B.sync();
    if (f[1])
        throw Error(SPH_ERR_UNSUPPORTED,
                    "more than two other cells within two stencil reaches share one cell's "
                    "spatial-hash bucket");
    if (f[0]) g.part = part.p;
}

// ---- single point-set helpers: standalone queries and static samples ----
static void cell_box(Batch &B, size_t cnt, double cs, const double *X, const double *Y,
                     const double *Z, int *ci, int *bbox) {
    hipLaunchKernelGGL(k_init_bbox, dim3(1), dim3(64), 0, B.s, bbox);
    HIPCK(hipGetLastError());
    if (cnt == 0) return;
    unsigned nbk = nblocks(cnt);
    B.ipart.reserve((size_t)nbk * 6);
    LAUNCH(k_cells, cnt, B.s, cnt, cs, X, Y, Z, ci, B.ipart.p);
    hipLaunchKernelGGL(k_fold_bbox, dim3(1), dim3(kBlock), 0, B.s, B.ipart.p, (int)nbk, bbox);
    HIPCK(hipGetLastError());
}

// counting sort by cell; members of a cell ordered by `ref` (or by index)
static void cell_sort(Batch &B, size_t cnt, const int *ci, const GridDev &g,
                      DBuf<uint32_t> &start, DBuf<uint32_t> &perm,
                      const uint32_t *ref = nullptr) {
    size_t ncell = (size_t)g.dim[0] * g.dim[1] * g.dim[2];
    B.ucnt.reserve(ncell);
    start.reserve(ncell + 1);
    B.ucursor.reserve(ncell);
    perm.reserve(std::max<size_t>(cnt, 1));
    LAUNCH(k_fill_u32, ncell, B.s, B.ucnt.p, ncell, 0u);
    LAUNCH(k_fill_u32, ncell, B.s, B.ucursor.p, ncell, 0u);
    LAUNCH(k_cell_count, cnt, B.s, cnt, ci, g, B.ucnt.p);
    scan_u32(B, B.ucnt.p, ncell, start.p);
    LAUNCH(k_cell_scatter, cnt, B.s, cnt, ci, g, start.p, B.ucursor.p, perm.p);
    LAUNCH(k_cell_order, ncell, B.s, ncell, start.p, perm.p, ref);
}

// ------------------------------------------------------------------------
// scene field storage
// ------------------------------------------------------------------------
static void reserve_scene(Scene &c, size_t n) {
    const size_t k = std::max<size_t>(n, 1);
    DBuf<double> *f64s[] = {&c.x,   &c.y,   &c.z,   &c.vx,  &c.vy,  &c.vz,  &c.ax,  &c.ay,
                            &c.az,  &c.m,   &c.V,   &c.p,   &c.s_f, &c.a_ff, &c.svx, &c.svy,
                            &c.svz, &c.sxp, &c.syp, &c.szp, &c.ix,  &c.iy,  &c.iz,  &c.ivx,
                            &c.ivy, &c.ivz, &c.tx,  &c.ty,  &c.tz,  &c.tvx, &c.tvy, &c.tvz,
                            &c.tm};
    for (auto *b : f64s) b->reserve(k);
    c.q4.reserve(4 * k);
    c.pa3.reserve(3 * k);
    c.sv4.reserve(4 * k);
    c.v3.reserve(3 * k);
    c.fid.reserve(k);
    c.ref.reserve(k);
    c.tfid.reserve(k);
    c.tref.reserve(k);
}

// StaticSampleBoundary samples at the current pose, sorted by (cell, index),
// with their grid: positions and velocities (update_positions_and_velocities,
// static_sample_boundary.rs:393-412, for samples on a rigid body: world
// positions from the body-frame ones and the rigid velocity there; static
// samples keep theirs and zero velocity) and, once their pseudo volumes are
// known, the volumes in sorted order.
static void pose_samples(Batch &B, Scene &c, Body &b) {
    const size_t m = b.ns;
    DBuf<double> hx, hy, hz, hv[3];
    std::vector<double> X(m), Y(m), Z(m), U[3];
    const bool moving = b.dynamic;
    if (moving)
        for (auto &u : U) u.resize(m);
    for (size_t k = 0; k < m; k++) {
        double w[3] = {b.spos[3 * k], b.spos[3 * k + 1], b.spos[3 * k + 2]};
        if (moving) {
            double p[3], u[3];
            b.rb.local_to_world(w, p);
            b.rb.velocity_at(p, u);
            for (int q = 0; q < 3; q++) {
                w[q] = p[q];
                U[q][k] = u[q];
            }
        }
        X[k] = w[0];
        Y[k] = w[1];
        Z[k] = w[2];
    }
    hx.upload(X.data(), m, B.s);
    hy.upload(Y.data(), m, B.s);
    hz.upload(Z.data(), m, B.s);
    if (moving)
        for (int q = 0; q < 3; q++) hv[q].upload(U[q].data(), m, B.s);
    B.sync();
    DBuf<int> sci;
    sci.reserve(3 * std::max<size_t>(m, 1));
    B.bbox.reserve(6);
    cell_box(B, m, c.P.kernel_support_radius, hx.p, hy.p, hz.p, sci.p, B.bbox.p);
    int bb[6];
    HIPCK(hipMemcpyAsync(bb, B.bbox.p, sizeof bb, hipMemcpyDeviceToHost, B.s));
    B.sync();
    if (m == 0)
        for (int k = 0; k < 6; k++) bb[k] = 0;
    b.sg = make_grid(bb, bb + 3);
    cell_sort(B, m, sci.p, b.sg, b.sgrid, b.sperm);
    b.sg.start = b.sgrid.p;
    b.sg.mkey = b.sperm.p;  // members in caller sample order (the bucket order)
    if (m) hash_partners(B, b.sg, b.spart, c.reach);
    const size_t m1 = std::max<size_t>(m, 1);
    b.sx.reserve(m1);
    b.sy.reserve(m1);
    b.sz.reserve(m1);
    b.svol.reserve(m1);
    LAUNCH(k_gather<double>, m, B.s, m, b.sperm.p, hx.p, b.sx.p);
    LAUNCH(k_gather<double>, m, B.s, m, b.sperm.p, hy.p, b.sy.p);
    LAUNCH(k_gather<double>, m, B.s, m, b.sperm.p, hz.p, b.sz.p);
    if (moving) {
        DBuf<double> *dst[3] = {&b.svx, &b.svy, &b.svz};
        for (int q = 0; q < 3; q++) {
            dst[q]->reserve(m1);
            LAUNCH(k_gather<double>, m, B.s, m, b.sperm.p, hv[q].p, dst[q]->p);
        }
        if (b.svol_set) LAUNCH(k_gather<double>, m, B.s, m, b.sperm.p, b.svol_c.p, b.svol.p);
    }
    B.sync();  // the host staging buffers go out of scope
}

// BoundaryType::initialize (static_sample_boundary.rs:354-390): the samples at
// their pose and their pseudo volumes from the boundary-boundary neighbors;
// volumes of samples on a rigid body are kept in caller order, since the
// samples are sorted again after every step of the body
static void prepare_samples(Batch &B, Scene &c, Body &b) {
    const size_t m = b.ns;
    pose_samples(B, c, b);
    if (b.dynamic && b.svol_set) return;  // volumes copied from the system it was cloned from
    DBuf<uint32_t> off, idx, scnt;
    off.reserve(m + 1);
    scnt.reserve(std::max<size_t>(m, 1));
    LAUNCH(k_neighbors<false>, m, B.s, m, b.sx.p, b.sy.p, b.sz.p, c.P.kernel_support_radius,
           c.reach, c.dthr, b.sg, b.sx.p, b.sy.p, b.sz.p, (const uint32_t *)nullptr, scnt.p,
           (uint32_t *)nullptr);
    uint32_t total = scan_u32(B, scnt.p, m, off.p);
    idx.reserve(std::max<uint32_t>(total, 1));
    LAUNCH(k_neighbors<true>, m, B.s, m, b.sx.p, b.sy.p, b.sz.p, c.P.kernel_support_radius,
           c.reach, c.dthr, b.sg, b.sx.p, b.sy.p, b.sz.p, off.p, (uint32_t *)nullptr, idx.p);
    LAUNCH(k_static_volume, m, B.s, m, c.K, c.P.boundary_rest_volume_weighting, b.sx.p, b.sy.p,
           b.sz.p, off.p, idx.p, b.svol.p);
    if (b.dynamic) {
        b.svol_c.reserve(std::max<size_t>(m, 1));
        LAUNCH(k_scatter<double>, m, B.s, m, b.sperm.p, b.svol.p, b.svol_c.p);
        b.svol_set = true;
    }
    B.sync();
}

static void init_boundaries(Batch &B, Scene &c) {
    if (c.boundaries_initialized) return;
    for (auto &b : c.bodies)
        if (b->kind == 0) prepare_samples(B, c, *b);
    c.boundaries_initialized = true;
}

// ------------------------------------------------------------------------
// batch layout
// ------------------------------------------------------------------------
static void upload_views(Batch &B) {
    std::vector<SceneView> v;
    v.reserve(B.sc.size());
    for (Scene *c : B.sc) v.push_back(c->view(B.cell_start.p));
    B.views.reserve(std::max<size_t>(v.size(), 1));
    B.put(B.views.p, v.data(), v.size() * sizeof(SceneView));
}

static void upload_bodies(Batch &B) {
    std::vector<BodyDev> v;
    for (Scene *c : B.sc)
        for (auto &b : c->bodies) v.push_back(b->view());
    B.bodies.reserve(std::max<size_t>(v.size(), 1));
    B.put(B.bodies.p, v.data(), v.size() * sizeof(BodyDev));
}

static BatchDev batch_dev(Batch &B) {
    BatchDev d{};
    d.views = B.views.p;
    d.pblk = B.pblk.p;
    d.cblk = B.cblk.p;
    d.ns = (int)B.sc.size();
    d.smap = nullptr;
    d.ns_all = (int)B.sc.size();
    d.ci = B.ci.p;
    d.perm = B.perm.p;
    d.cnt = B.cnt.p;
    d.slice_len = B.slice_len.p;
    d.slice_off = B.slice_off.p;
    d.ell_j = B.ell_j.p;
    d.ell_gx = B.ell_gx.p;
    d.ell_gy = B.ell_gy.p;
    d.ell_gz = B.ell_gz.p;
    d.cell_cnt = B.cell_cnt.p;
    d.cell_start = B.cell_start.p;
    d.cursor = B.cursor.p;
    d.bodies = B.bodies.p;
    d.partials = B.partials.p;
    d.active = B.d_active();
    d.iters = B.d_iters();
    d.chunk_err = B.chunk_err();
    d.chunk_ran = B.chunk_ran();
    return d;
}

// particle-side placement of every scene; per-particle batch buffers sized
static void layout(Batch &B) {
    const size_t ns = B.sc.size();
    B.hp_blk.assign(ns + 1, 0);
    size_t parts = 0, slices = 0, bodies = 0, blocks = 0;
    for (size_t s = 0; s < ns; s++) {
        Scene &c = *B.sc[s];
        if (c.bodies.size() > (size_t)kMaxBodiesPerScene)
            throw Error(SPH_ERR_UNSUPPORTED, "too many boundaries in one system");
        c.part_base = (uint32_t)parts;
        c.slice_base = (uint32_t)slices;
        c.body_base = (uint32_t)bodies;
        B.hp_blk[s] = (uint32_t)blocks;
        parts += c.n;
        slices += (c.n + 63) / 64;
        bodies += c.bodies.size();
        blocks += nblocks(c.n);
    }
    B.hp_blk[ns] = (uint32_t)blocks;
    if (parts >= 0xffffffffull) throw Error(SPH_ERR_UNSUPPORTED, "more than 2^32-1 particles");
    B.total_particles = parts;
    B.total_pblocks = blocks;
    B.total_slices = slices;
    B.pblk.reserve(ns + 1);
    B.put(B.pblk.p, B.hp_blk.data(), (ns + 1) * sizeof(uint32_t));
    B.ci.reserve(3 * std::max<size_t>(parts, 1));
    B.perm.reserve(std::max<size_t>(parts, 1));
    B.cnt.reserve(std::max<size_t>(parts, 1));
    B.slice_len.reserve(slices + 1);
    B.slice_off.reserve(slices + 1);
    B.partials.reserve((blocks + 1) * kMaxBodiesPerScene * 6 + 64);
    B.red.reserve(std::max<size_t>(ns, 1) * kMaxBodiesPerScene * 6 + 64);
    B.ensure_pinned(std::max<size_t>(ns, 1) * (kMaxBodiesPerScene * 6 + 2 * kMaxChunk + 2) + 64);
    B.solve_io.reserve(std::max<size_t>(ns, 1) * (kMaxChunk + (kMaxChunk + 2) / 2 + 1));
    upload_bodies(B);
    upload_views(B);
}

// ------------------------------------------------------------------------
// phases of System::update over all scenes of a batch
// ------------------------------------------------------------------------
// max |v| and the below-threshold count of every scene
static void speed_below(Batch &B, std::vector<double> &vmax, std::vector<double> &below) {
    const size_t ns = B.sc.size();
    vmax.assign(ns, 0.0);
    below.assign(ns, 0.0);
    if (B.total_pblocks == 0) return;
    BatchDev d = batch_dev(B);
    PLAUNCH(B, SPH_PROF_KERNELS, 0, kb_speed_below, B.total_pblocks, d);
    hipLaunchKernelGGL(kb_fold_speed, dim3((unsigned)ns), dim3(kBlock), 0, B.s, B.partials.p,
                       B.pblk.p, (int)ns, B.red.p);
    HIPCK(hipGetLastError());
    HIPCK(hipMemcpyAsync(B.pinned, B.red.p, 2 * ns * sizeof(double), hipMemcpyDeviceToHost, B.s));
    B.sync();
    for (size_t s = 0; s < ns; s++) {
        vmax[s] = B.pinned[2 * s];
        below[s] = B.pinned[2 * s + 1];
    }
}

// Fluid::disable/swap/drop_inactive on the reference index order (fluid.rs:159-210)
static void compact_below(Batch &B, Scene &c) {
    const size_t n = c.n;
    std::vector<double> zh(n);
    std::vector<uint32_t> rh(n);
    c.z.download(zh.data(), n, B.s);
    c.ref.download(rh.data(), n, B.s);
    B.sync();
    std::vector<uint32_t> order(n);  // order[r] = storage slot of reference index r
    for (size_t k = 0; k < n; k++) order[rh[k]] = (uint32_t)k;
    size_t id = 0, nn = n;
    while (id < nn) {
        if (zh[order[id]] < c.P.disable_particles_below) {
            nn -= 1;
            std::swap(order[id], order[nn]);
        } else {
            id++;
        }
    }
    std::vector<uint32_t> newref(n, 0xffffffffu);
    for (size_t r = 0; r < nn; r++) newref[order[r]] = (uint32_t)r;
    std::vector<uint32_t> keep, kref;
    keep.reserve(nn);
    kref.reserve(nn);
    for (size_t k = 0; k < n; k++)
        if (newref[k] != 0xffffffffu) {
            keep.push_back((uint32_t)k);
            kref.push_back(newref[k]);
        }
    DBuf<uint32_t> sel;
    sel.upload(keep.data(), nn, B.s);
    LAUNCH(k_gather<double>, nn, B.s, nn, sel.p, c.x.p, c.tx.p);
    LAUNCH(k_gather<double>, nn, B.s, nn, sel.p, c.y.p, c.ty.p);
    LAUNCH(k_gather<double>, nn, B.s, nn, sel.p, c.z.p, c.tz.p);
    LAUNCH(k_gather<double>, nn, B.s, nn, sel.p, c.vx.p, c.tvx.p);
    LAUNCH(k_gather<double>, nn, B.s, nn, sel.p, c.vy.p, c.tvy.p);
    LAUNCH(k_gather<double>, nn, B.s, nn, sel.p, c.vz.p, c.tvz.p);
    LAUNCH(k_gather<double>, nn, B.s, nn, sel.p, c.m.p, c.tm.p);
    LAUNCH(k_gather<uint32_t>, nn, B.s, nn, sel.p, c.fid.p, c.tfid.p);
    const bool slot = c.P.integrator == SPH_INTEGRATOR_VERLET;  // the slot persists
    if (slot) {
        LAUNCH(k_gather<double>, nn, B.s, nn, sel.p, c.ix.p, c.sxp.p);
        LAUNCH(k_gather<double>, nn, B.s, nn, sel.p, c.iy.p, c.syp.p);
        LAUNCH(k_gather<double>, nn, B.s, nn, sel.p, c.iz.p, c.szp.p);
    }
    c.tref.upload(kref.data(), nn, B.s);
    B.sync();
    if (slot) {
        c.ix.swap(c.sxp);
        c.iy.swap(c.syp);
        c.iz.swap(c.szp);
    }
    c.x.swap(c.tx);
    c.y.swap(c.ty);
    c.z.swap(c.tz);
    c.vx.swap(c.tvx);
    c.vy.swap(c.tvy);
    c.vz.swap(c.tvz);
    c.m.swap(c.tm);
    c.fid.swap(c.tfid);
    c.ref.swap(c.tref);
    c.n = nn;
}

// cell boxes, hash guard, grids and the batch cell table
static void cells_and_grids(Batch &B) {
    const size_t ns = B.sc.size();
    std::vector<int> boxes(6 * ns, 0);
    if (B.total_pblocks) {
        BatchDev d = batch_dev(B);
        PLAUNCH(B, SPH_PROF_KERNELS, 0, kb_cells, B.total_pblocks, d);
        B.ipart.reserve(6 * ns + 6);
        hipLaunchKernelGGL(kb_fold_bbox, dim3((unsigned)ns), dim3(kBlock), 0, B.s,
                           reinterpret_cast<const int *>(B.partials.p), B.pblk.p, (int)ns,
                           B.ipart.p);
        HIPCK(hipGetLastError());
        int *hb = reinterpret_cast<int *>(B.pinned);
        HIPCK(hipMemcpyAsync(hb, B.ipart.p, 6 * ns * sizeof(int), hipMemcpyDeviceToHost, B.s));
        B.sync();
        memcpy(boxes.data(), hb, 6 * ns * sizeof(int));
    }
    B.hc_blk.assign(ns + 1, 0);
    size_t cells = 0, cblocks = 0;
    for (size_t s = 0; s < ns; s++) {
        Scene &c = *B.sc[s];
        int *bb = &boxes[6 * s];
        if (c.n == 0)
            for (int k = 0; k < 6; k++) bb[k] = 0;
        memcpy(c.box, bb, sizeof c.box);
        const bool shared = c.n && hash_guard(B, c.checked_box, c.hash_box_valid,
                                              c.hash_collides, bb, bb + 3, c.reach);
        c.grid = make_grid(bb, bb + 3);
        if (shared) hash_partners(B, c.grid, c.hpart, c.reach);
        c.cell_base = (uint32_t)cells;
        B.hc_blk[s] = (uint32_t)cblocks;
        cells += c.ncells();
        cblocks += nblocks(c.ncells());
    }
    B.hc_blk[ns] = (uint32_t)cblocks;
    if (cells >= 0xffffffffull) throw Error(SPH_ERR_UNSUPPORTED, "cell table too large");
    B.total_cells = cells;
    B.cblk.reserve(ns + 1);
    B.put(B.cblk.p, B.hc_blk.data(), (ns + 1) * sizeof(uint32_t));
    B.cell_cnt.reserve(cells + 1);
    B.cell_start.reserve(cells + 1);
    B.cursor.reserve(cells + 1);
    upload_views(B);
}

// counting sort of every scene by (cell, reference index) and the permutation
// of the persistent fields (System::update -> N::find_samples, sph/mod.rs:701-706)
static void sort_scenes(Batch &B) {
    BatchDev d = batch_dev(B);
    LAUNCH(k_fill_u32, B.total_cells, B.s, B.cell_cnt.p, B.total_cells, 0u);
    LAUNCH(k_fill_u32, B.total_cells, B.s, B.cursor.p, B.total_cells, 0u);
    PLAUNCH(B, SPH_PROF_KERNELS, 0, kb_cell_count, B.total_pblocks, d);
    // every particle of the batch lies in one cell: the total is known
    scan_u32(B, B.cell_cnt.p, B.total_cells, B.cell_start.p, (int64_t)B.total_particles);
    PLAUNCH(B, SPH_PROF_KERNELS, 0, kb_cell_scatter, B.total_pblocks, d);
    PLAUNCH(B, SPH_PROF_KERNELS, 0, kb_cell_order, B.hc_blk.back(), d);
    PLAUNCH(B, SPH_PROF_KERNELS, 0, kb_permute, B.total_pblocks, d);
    for (Scene *c : B.sc) {
        if (c->P.integrator == SPH_INTEGRATOR_VERLET) {  // kb_permute moved the slot there
            c->ix.swap(c->sxp);
            c->iy.swap(c->syp);
            c->iz.swap(c->szp);
        }
        c->x.swap(c->tx);
        c->y.swap(c->ty);
        c->z.swap(c->tz);
        c->vx.swap(c->tvx);
        c->vy.swap(c->tvy);
        c->vz.swap(c->tvz);
        c->m.swap(c->tm);
        c->fid.swap(c->tfid);
        c->ref.swap(c->tref);
    }
    upload_views(B);
}

// neighbor counts, slice widths and the pair cache's size
static void neighbor_counts(Batch &B) {
    BatchDev d = batch_dev(B);
    PLAUNCH(B, SPH_PROF_NEIGHBOR_COUNT, B.total_particles, kb_count_ell, B.total_pblocks, d);
    uint32_t rows = scan_u32(B, B.slice_len.p, B.total_slices, B.slice_off.p);
    uint64_t slots = (uint64_t)rows * 64;
    if (slots > 0xffffffffull) throw Error(SPH_ERR_UNSUPPORTED, "neighbor cache exceeds 2^32 slots");
    B.ell_j.reserve(std::max<uint64_t>(slots, 1));
    if (B.grad_cache) {
        B.ell_gx.reserve(std::max<uint64_t>(slots, 1));
        B.ell_gy.reserve(std::max<uint64_t>(slots, 1));
        B.ell_gz.reserve(std::max<uint64_t>(slots, 1));
    }
}

// B::find_boundary_samples (sph/mod.rs:707-712)
static void boundary_samples(Batch &B) {
    bool any_vm = false;
    for (Scene *c : B.sc) {
        const size_t n = c->n;
        for (auto &bp : c->bodies) {
            Body &b = *bp;
            if (b.kind == 0) {
                // per-scene CSR over the static samples
                b.soff.reserve(n + 1);
                b.scnt.reserve(std::max<size_t>(n, 1));
                // shared buckets among the sample cells were found at set-up (b.sg.part)
                const double range = c->P.kernel_support_radius;
                LAUNCH(k_neighbors<false>, n, B.s, n, c->x.p, c->y.p, c->z.p, range, c->reach,
                       c->dthr, b.sg, b.sx.p, b.sy.p, b.sz.p, (const uint32_t *)nullptr, b.scnt.p,
                       (uint32_t *)nullptr);
                uint32_t total = scan_u32(B, b.scnt.p, n, b.soff.p);
                b.sidx.reserve(std::max<uint32_t>(total, 1));
                LAUNCH(k_neighbors<true>, n, B.s, n, c->x.p, c->y.p, c->z.p, range, c->reach,
                       c->dthr, b.sg, b.sx.p, b.sy.p, b.sz.p, b.soff.p, (uint32_t *)nullptr,
                       b.sidx.p);
            } else {
                any_vm = true;
                b.flag.reserve(std::max<size_t>(n, 1));
                DBuf<double> *f[] = {&b.bx, &b.by, &b.bz, &b.bvx, &b.bvy, &b.bvz, &b.bvol};
                for (auto *q : f) q->reserve(std::max<size_t>(n, 1));
            }
        }
    }
    upload_bodies(B);
    if (any_vm) {
        BatchDev d = batch_dev(B);
        PLAUNCH(B, SPH_PROF_BOUNDARY_SAMPLE, B.total_particles, kb_vm_sample, B.total_pblocks, d);
    }
}

// neighbor indices and the volume in one sweep (kb_sweep_ell) into a pair
// cache of fixed width: the largest neighbor count of the last sweep plus 8,
// at least 16. A sweep that finds more neighbors than fit runs again with a
// wider cache. Boundary samples must be current (the volume reads them).
static void sweep_neighbors(Batch &B) {
    if (B.total_particles == 0) return;
    B.max_cnt.reserve(1);
    uint32_t cap = B.ell_cap ? B.ell_cap : 64;
    for (;;) {
        const uint64_t slots = (uint64_t)B.total_slices * 64 * cap;
        if (slots > 0xffffffffull) throw Error(SPH_ERR_UNSUPPORTED, "neighbor cache exceeds 2^32 slots");
        B.ell_j.reserve(std::max<uint64_t>(slots, 1));
        LAUNCH(kb_fixed_slices, B.total_slices + 1, B.s, B.slice_off.p, B.total_slices, cap);
        LAUNCH(k_fill_u32, 1, B.s, B.max_cnt.p, (size_t)1, 0u);
        BatchDev d = batch_dev(B);
        PLAUNCH(B, SPH_PROF_NEIGHBOR_COUNT, B.total_particles, kb_sweep_ell, B.total_pblocks, d,
                cap, B.max_cnt.p);
        uint32_t mx = 0;
        HIPCK(hipMemcpyAsync(B.pinned, B.max_cnt.p, sizeof mx, hipMemcpyDeviceToHost, B.s));
        B.sync();
        memcpy(&mx, B.pinned, sizeof mx);
        const uint32_t want = std::max<uint32_t>(16, ((mx + 8 + 7) / 8) * 8);
        if (mx <= cap) {
            B.ell_cap = want;
            return;
        }
        cap = want;  // overflow: the cache was too narrow for this sweep
    }
}

// neighbor indices, then grad W into the pair cache and the volume (get_volume)
static void pairs_and_volume(Batch &B) {
    BatchDev d = batch_dev(B);
    PLAUNCH(B, SPH_PROF_PAIRS_VOLUME, B.total_particles, kb_fill_ell, B.total_pblocks, d);
    pair_pass(B, PP_VOLUME, SPH_PROF_PAIRS_VOLUME, k_volume_tab, d);
}

static bool any_dynamic(const Batch &B) {
    for (const Scene *c : B.sc)
        for (auto &b : c->bodies)
            if (b->dynamic && c->n) return true;
    return false;
}

// reduce per-workgroup body partials per scene and add them to the rigid bodies
static void apply_body_partials(Batch &B) {
    if (!any_dynamic(B) || B.total_pblocks == 0) return;
    const size_t ns = B.sc.size();
    const int rows = kMaxBodiesPerScene * 6;
    hipLaunchKernelGGL(kb_fold_sum, dim3((unsigned)(ns * rows)), dim3(kBlock), 0, B.s,
                       B.partials.p, B.pblk.p, (int)ns, rows, B.red.p);
    HIPCK(hipGetLastError());
    HIPCK(hipMemcpyAsync(B.pinned, B.red.p, ns * rows * sizeof(double), hipMemcpyDeviceToHost, B.s));
    B.sync();
    for (size_t s = 0; s < ns; s++) {
        Scene &c = *B.sc[s];
        if (c.n == 0) continue;
        for (size_t bi = 0; bi < c.bodies.size(); bi++) {
            Body &b = *c.bodies[bi];
            if (!b.dynamic) continue;
            const double *f = B.pinned + s * rows + bi * 6;
            b.rb.add_force_torque(f);
            for (int k = 0; k < 6; k++) b.applied[k] += f[k];
        }
    }
}

static void pressure_accel(Batch &B, int overwrite) {
    BatchDev d = batch_dev(B);
    pair_pass(B, PP_PRESSURE_ACCEL, SPH_PROF_PRESSURE_ACCEL, k_pressure_accel_tab, d, 0,
              overwrite, 1);
    apply_body_partials(B);
}

// `chunk` Jacobi iterations of every active scene: pressure acceleration into
// the scratch, the Jacobi update, and the per-scene fold + termination test
// (d covers the scenes still iterating: d.ns of them, `grid` workgroups)
static void jacobi_chunk_launches(Batch &B, const BatchDev &d, size_t grid, int after,
                                  int chunk) {
    for (int k = 0; k < chunk; k++) {
        if (B.tuner.measuring())
            B.variant[PP_PRESSURE_ACCEL] = B.variant[PP_JACOBI] =
                (int)(B.tuner.rotation++ % kPairVariants);
        pair_pass_grid(B, grid, PP_PRESSURE_ACCEL, SPH_PROF_PRESSURE_ACCEL, k_pressure_accel_tab,
                       d, 1, 1, 0);
        pair_pass_grid(B, grid, PP_JACOBI, SPH_PROF_JACOBI, k_jacobi_tab, d);
        hipLaunchKernelGGL(kb_fold_converge, dim3((unsigned)d.ns), dim3(kBlock), 0, B.s, d,
                           after, k);
        HIPCK(hipGetLastError());
    }
}

// The chunk's launches in the form a graph replays: the kernels read the batch
// descriptor from B.desc (uploaded before each chunk) and run over grids
// rounded up to powers of two (gp workgroups for the pair passes, gs for the
// per-scene fold), so their arguments do not depend on the layout.
typedef void (*KernDIII)(const BatchDev *, int, int, int);
typedef void (*KernD)(const BatchDev *);
#define PAD_F(U, W, P) &kb_pressure_accel_d<false, U, W, P>,
#define PAD_T(U, W, P) &kb_pressure_accel_d<true, U, W, P>,
#define JAD_F(U, W, P) &kb_jacobi_d<false, U, W, P>,
#define JAD_T(U, W, P) &kb_jacobi_d<true, U, W, P>,
static const KernDIII k_pressure_accel_d_tab[2][kPairVariants] = {{SPH_PAIR_VARIANTS(PAD_F)},
                                                                  {SPH_PAIR_VARIANTS(PAD_T)}};
static const KernD k_jacobi_d_tab[2][kPairVariants] = {{SPH_PAIR_VARIANTS(JAD_F)},
                                                       {SPH_PAIR_VARIANTS(JAD_T)}};

static unsigned grid_bucket(size_t n) {
    unsigned g = 8;
    while (g < n) g <<= 1;
    return g;
}

static void jacobi_chunk_desc_launches(Batch &B, unsigned gp, unsigned gs, int after, int chunk) {
    const int gc = B.grad_cache ? 1 : 0;
    const KernDIII pa = k_pressure_accel_d_tab[gc][B.variant[PP_PRESSURE_ACCEL]];
    const KernD ja = k_jacobi_d_tab[gc][B.variant[PP_JACOBI]];
    const BatchDev *dd = B.desc.p;
    for (int k = 0; k < chunk; k++) {
        hipLaunchKernelGGL(pa, dim3(gp), dim3(kBlock), 0, B.s, dd, 1, 1, 0);
        HIPCK(hipGetLastError());
        hipLaunchKernelGGL(ja, dim3(gp), dim3(kBlock), 0, B.s, dd);
        HIPCK(hipGetLastError());
        hipLaunchKernelGGL(kb_fold_converge_d, dim3(gs), dim3(kBlock), 0, B.s, dd, after, k);
        HIPCK(hipGetLastError());
    }
}

// A chunk is replayed from a HIP graph of its descriptor-form launches: one
// submission instead of three per iteration. The graph's arguments are the
// descriptor's address, the termination rule, the chunk length, the grid
// buckets and the code shapes, so a world reuses its graphs across
// admissions and departures. A key is captured on its second use; timing and
// tuning runs launch directly. Executable graphs live until the batch ends or
// graphs are switched off (none is destroyed while stepping); once
// kMaxGraphs exist, new keys launch directly.
static void jacobi_chunk(Batch &B, const BatchDev &d, size_t grid, int after, int chunk) {
    if (!B.use_graphs || B.prof.on || B.tuner.measuring()) {
        jacobi_chunk_launches(B, d, grid, after, chunk);
        return;
    }
    const unsigned gp = grid_bucket(kXcds * ((grid + kXcds - 1) / kXcds));
    const unsigned gs = grid_bucket((size_t)d.ns);
    const int meta[7] = {after, chunk, (int)gp, (int)gs, B.grad_cache ? 1 : 0,
                         B.variant[PP_PRESSURE_ACCEL], B.variant[PP_JACOBI]};
    const std::string key(reinterpret_cast<const char *>(meta), sizeof meta);
    hipGraphExec_t exec = nullptr;
    for (const Batch::ChunkGraph &g : B.graphs)
        if (g.key == key) exec = g.exec;
    const bool capture = !exec && B.graphs.size() < kMaxGraphs &&
                         std::find(B.seen.begin(), B.seen.end(), key) != B.seen.end();
    if (!exec && !capture) {
        if (std::find(B.seen.begin(), B.seen.end(), key) == B.seen.end()) {
            B.seen.push_back(key);
            if (B.seen.size() > 16) B.seen.erase(B.seen.begin());
        }
        jacobi_chunk_launches(B, d, grid, after, chunk);
        return;
    }
    // the descriptor the replayed kernels read, ahead of them in stream order
    B.desc.reserve(1);
    B.put(B.desc.p, &d, sizeof d);
    if (capture) {
        hipGraph_t graph = nullptr;
        HIPCK(hipStreamBeginCapture(B.s, hipStreamCaptureModeThreadLocal));
        try {
            jacobi_chunk_desc_launches(B, gp, gs, after, chunk);
        } catch (...) {
            (void)hipStreamEndCapture(B.s, &graph);
            if (graph) (void)hipGraphDestroy(graph);
            throw;
        }
        HIPCK(hipStreamEndCapture(B.s, &graph));
        hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        HIPCK(e);
        B.graphs.push_back({key, exec});
    }
    HIPCK(hipGraphLaunch(exec, B.s));
}

// IISPH::resolve_pressure_globally (iisph.rs:378-560) for every scene at once;
// after < 0: target density error, else AfterIteration(after). Iterations are
// launched in chunks; after each iteration kb_fold_converge applies every
// scene's own termination test on the device, and the iterations of finished
// scenes return at once, so a chunk may run past a scene's last iteration
// without changing it. The host waits once per chunk: the first chunk covers
// the iterations the solve needs at least (after + 1, or 3 and the count of
// the last target-terminated solve), later chunks two each.
static void resolve_pressure(Batch &B, int vp, int after) {
    const size_t ns = B.sc.size();
    BatchDev d = batch_dev(B);
    pair_pass(B, PP_SOURCE_DIAG, SPH_PROF_SOURCE_DIAG, k_source_diag_tab, d, vp);
    std::vector<uint32_t> act(ns), zero(ns, 0u);
    bool any = false;
    for (size_t s = 0; s < ns; s++) {
        Scene &c = *B.sc[s];
        c.iter = 0;
        c.err = INFINITY;
        c.err_hist.clear();
        c.active = c.n > 0;
        act[s] = c.active ? 1u : 0u;
        any = any || c.active;
    }
    act.insert(act.end(), zero.begin(), zero.end());  // active flags, then iteration counts
    B.put(B.d_active(), act.data(), 2 * ns * sizeof(uint32_t));
    int chunk = after >= 0 ? after + 1 : std::max<int>(3, (int)B.iter_hint);
    uint32_t most = 0;
    std::vector<uint32_t> ablk, amap;
    while (any) {
        chunk = std::min(std::max(chunk, 1), kMaxChunk);
        if (B.tuner.measuring()) chunk = 1;  // every timed iteration does work
        for (Scene *c : B.sc)  // no scene iterates past its cap
            if (c->active) chunk = std::min<int>(chunk, (int)(c->max_iter - c->iter));
        // launch over the scenes still iterating only
        ablk.assign(1, 0u);
        amap.clear();
        for (size_t s = 0; s < ns; s++)
            if (B.sc[s]->active) {
                amap.push_back((uint32_t)s);
                ablk.push_back(ablk.back() + (B.hp_blk[s + 1] - B.hp_blk[s]));
            }
        BatchDev da = d;
        if (amap.size() < ns) {
            B.ablk.reserve(ns + 1);
            B.amap.reserve(ns);
            B.put(B.ablk.p, ablk.data(), ablk.size() * sizeof(uint32_t));
            B.put(B.amap.p, amap.data(), amap.size() * sizeof(uint32_t));
            da.pblk = B.ablk.p;
            da.smap = B.amap.p;
            da.ns = (int)amap.size();
        }
        jacobi_chunk(B, da, ablk.back(), after, chunk);
        // errors, ran flags and active flags in one copy, in the device layout
        double *errs = B.pinned;
        uint32_t *ran = reinterpret_cast<uint32_t *>(errs + (size_t)kMaxChunk * ns);
        uint32_t *alive = ran + (size_t)kMaxChunk * ns;
        HIPCK(hipMemcpyAsync(errs, B.solve_io.p,
                             (size_t)kMaxChunk * ns * (sizeof(double) + sizeof(uint32_t)) +
                                 ns * sizeof(uint32_t),
                             hipMemcpyDeviceToHost, B.s));
        B.sync();
        if (B.prof.on) B.prof.drain();
        any = false;
        for (size_t s = 0; s < ns; s++) {
            Scene &c = *B.sc[s];
            // a scene outside this chunk's launch left the loop earlier: its
            // ran and error slots still hold the chunk it last ran in
            if (!c.active) {
                most = std::max(most, c.iter);
                continue;
            }
            for (int k = 0; k < chunk; k++)
                if (ran[(size_t)k * ns + s]) {
                    c.err = errs[(size_t)k * ns + s];
                    c.err_hist.push_back(c.err);
                    c.iter++;
                }
            c.active = alive[s] != 0;
            most = std::max(most, c.iter);
            if (c.active && c.iter >= c.max_iter) {
                // this scene fails; the others keep iterating
                c.failed = SPH_ERR_UNSUPPORTED;
                c.fail_msg = "pressure solve did not converge within " +
                             std::to_string(c.max_iter) + " iterations";
                c.active = false;
                const uint32_t zero = 0;
                B.put(B.d_active() + s, &zero, sizeof zero);
            }
            any = any || c.active;
        }
        chunk = 2;
    }
    if (after < 0 && most > 0) B.iter_hint = most;
    for (Scene *c : B.sc) {
        c->last_iter = c->iter;
        c->last_err = c->n ? c->err : 0.0;
        c->active = false;
    }
}

// PressureSolver::solve_and_add_acceleration for every scene (one solver per batch)
static void solve_pressure(Batch &B) {
    if (B.sc.empty()) return;
    const int solver = B.sc[0]->P.solver;
    BatchDev d = batch_dev(B);
    if (solver == SPH_SOLVER_SESPH) {  // sesph.rs:31-67
        PLAUNCH(B, SPH_PROF_KERNELS, 0, kb_sesph_pressure, B.total_pblocks, d);
        pressure_accel(B, 0);
    } else if (solver == SPH_SOLVER_SESPH_SPLITTING) {  // sesph_with_splitting.rs:31-81
        PLAUNCH(B, SPH_PROF_KERNELS, 0, kb_pred_vel, B.total_pblocks, d, 0);
        PLAUNCH(B, SPH_PROF_KERNELS, 0, kb_sesph_split_density, B.total_pblocks, d);
        pressure_accel(B, 0);
    } else if (solver == SPH_SOLVER_IISPH) {  // iisph.rs:55-96
        PLAUNCH(B, SPH_PROF_KERNELS, 0, kb_pred_vel, B.total_pblocks, d, 0);
        resolve_pressure(B, 1, -1);
        pressure_accel(B, 0);
    } else {  // iisph_optimized_source_term.rs:53-217
        PLAUNCH(B, SPH_PROF_KERNELS, 0, kb_pred_vel, B.total_pblocks, d, 0);
        resolve_pressure(B, 0, 3);
        for (Scene *c : B.sc) c->eqs1_iter = c->last_iter;
        pressure_accel(B, 1);
        PLAUNCH(B, SPH_PROF_KERNELS, 0, kb_pred_vel, B.total_pblocks, d, 1);
        PLAUNCH(B, SPH_PROF_KERNELS, 0, kb_pred_pos, B.total_pblocks, d);
        resolve_pressure(B, 1, -1);
        pressure_accel(B, 1);
        PLAUNCH(B, SPH_PROF_KERNELS, 0, kb_ost_commit, B.total_pblocks, d);
        // solver slots become the integrator slots (TakePredicted commits them)
        for (Scene *c : B.sc) {
            c->ix.swap(c->sxp);
            c->iy.swap(c->syp);
            c->iz.swap(c->szp);
            c->ivx.swap(c->svx);
            c->ivy.swap(c->svy);
            c->ivz.swap(c->svz);
        }
        upload_views(B);
    }
}

// reset_acceleration + gravity + viscosity (sph/mod.rs:640-657)
static void nonpressure(Batch &B) {
    // gravity onto dynamic bodies first (non_pressure_accelerations.rs:64-103)
    for (Scene *c : B.sc) {
        for (auto &bp : c->bodies) {
            Body &b = *bp;
            if (!b.dynamic) continue;
            double acc[3];
            if (c->P.gravity_mode == SPH_GRAVITY_UNIFORM) {
                acc[0] = 0.0;
                acc[1] = 0.0;
                acc[2] = -9.81;
            } else {
                double dd[3] = {c->P.gravity_center[0] - b.rb.com[0],
                                c->P.gravity_center[1] - b.rb.com[1],
                                c->P.gravity_center[2] - b.rb.com[2]};
                double dist = std::sqrt(((dd[0] * dd[0] + dd[1] * dd[1]) + dd[2] * dd[2]) + 1e-4);
                for (int k = 0; k < 3; k++) acc[k] = (9.81 * dd[k]) / dist;
            }
            double f[3] = {b.rb.mass * acc[0], b.rb.mass * acc[1], b.rb.mass * acc[2]};
            b.rb.add_force(f, b.rb.com);
            for (int k = 0; k < 3; k++) b.applied[k] += f[k];
        }
    }
    BatchDev d = batch_dev(B);
    pair_pass(B, PP_NONPRESSURE, SPH_PROF_NONPRESSURE, k_nonpressure_tab, d);
    apply_body_partials(B);
}

static void measure_sums(Batch &B, std::vector<double> &out) {
    const size_t ns = B.sc.size();
    out.assign(3 * ns, 0.0);
    if (B.total_pblocks == 0) return;
    BatchDev d = batch_dev(B);
    PLAUNCH(B, SPH_PROF_KERNELS, 0, kb_measure, B.total_pblocks, d);
    hipLaunchKernelGGL(kb_fold_sum, dim3((unsigned)(3 * ns)), dim3(kBlock), 0, B.s, B.partials.p,
                       B.pblk.p, (int)ns, 3, B.red.p);
    HIPCK(hipGetLastError());
    HIPCK(hipMemcpyAsync(B.pinned, B.red.p, 3 * ns * sizeof(double), hipMemcpyDeviceToHost, B.s));
    B.sync();
    for (size_t k = 0; k < 3 * ns; k++) out[k] = B.pinned[k];
}

// neighbors + boundary samples + pair cache + volume for the batch's scenes
static void prepare(Batch &B) {
    cells_and_grids(B);
    sort_scenes(B);
    if (B.sweep && !B.grad_cache) {
        boundary_samples(B);
        sweep_neighbors(B);
    } else {
        neighbor_counts(B);
        boundary_samples(B);
        pairs_and_volume(B);
    }
}

// System::update (sph/mod.rs:674-731) for every scene of the batch
// the first failure of a scene in the last step, or nullptr
static const Scene *first_failed(const Batch &B) {
    for (const Scene *c : B.sc)
        if (c->failed) return c;
    return nullptr;
}

static void update(Batch &B) {
    for (Scene *c : B.sc) {
        c->failed = 0;
        c->fail_msg.clear();
    }
    layout(B);
    std::vector<double> vmax, below;
    speed_below(B, vmax, below);
    bool compacted = false;
    for (size_t s = 0; s < B.sc.size(); s++)
        if (below[s] > 0.0) {
            compact_below(B, *B.sc[s]);
            compacted = true;
        }
    if (compacted) {
        layout(B);
        speed_below(B, vmax, below);  // max speed over the remaining particles
    }
    for (size_t s = 0; s < B.sc.size(); s++) {
        Scene &c = *B.sc[s];
        if (c.P.cfl) {  // set_cfl_time_step (sph/mod.rs:825-829)
            double cfl = c.P.cfl_number * c.P.rest_density_grid_spacing / vmax[s];
            c.dt = std::fmin(c.P.max_time_increment, cfl);
        }
        for (auto &b : c.bodies)
            for (int k = 0; k < 6; k++) b->applied[k] = 0;
    }
    prepare(B);
    nonpressure(B);
    solve_pressure(B);
    std::vector<double> sums;
    measure_sums(B, sums);
    for (size_t s = 0; s < B.sc.size(); s++) {
        Scene &c = *B.sc[s];
        c.avg_density = c.n ? sums[3 * s] / (double)c.n : 0.0;
    }
    if (B.prof.on) {
        B.sync();
        B.prof.drain();
    }
}

// ------------------------------------------------------------------------
// rigid-rigid contact (VolumeMapBoundary::step_forward_in_time,
// volume_map_boundary.rs:1440-1453): the broad phase and sub-cell grid on the
// host, the contact cells' quadrature on the device (k_contact_cells), then
// compute_diagonal's closing terms and the NNCG solve (solve_for_p_c,
// :186-637) over the active cells on the host, in binary64 with the
// reference's operation order
// ------------------------------------------------------------------------
// a contact-cell record: body a, body b, volume, center[3], f_a[3], tau_a[3],
// f_b[3], tau_b[3], a_cc pressure, a_cc friction, p, lambda[3]
constexpr int kContactRec = 24;
namespace contact {
constexpr double kCellSizeFactor = 0.5;  // CONTACT_CELL_SIZE_FACTOR
constexpr double kFriction = 0.5;        // FRICTION_COEFFICIENT
constexpr double kRelaxation = 0.5;      // RELAXATION
constexpr double kMinDiagonal = 1e-35;   // MIN_DIAGONAL_ELEMENT
constexpr double kFbPressure = 1e-6, kFbFriction = 1e-3;
// the reference iterates until converged (warning past 100 iterations)
constexpr uint32_t kMaxIterations = 100000;
enum { A = 0, B = 1, VOL = 2, CEN = 3, FA = 6, TA = 9, FB = 12, TB = 15, ACCP = 18, ACCF = 19,
       P = 20, LAM = 21 };

static inline double dot(const double *a, const double *b) {
    return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2];
}
static inline double norm(const double *a) { return std::sqrt(dot(a, a)); }
static inline void cross(const double *a, const double *b, double *o) {
    const double x = a[1] * b[2] - a[2] * b[1];
    const double y = a[2] * b[0] - a[0] * b[2];
    const double z = a[0] * b[1] - a[1] * b[0];
    o[0] = x;
    o[1] = y;
    o[2] = z;
}
static inline void sub(const double *a, const double *b, double *o) {
    for (int k = 0; k < 3; k++) o[k] = a[k] - b[k];
}
// UnitQuaternion * Vector3: t = 2 (q_v x v); v' = t w + q_v x t + v
static void rotate(const double q[4], const double v[3], double o[3]) {
    const double qv[3] = {q[1], q[2], q[3]};
    double t[3], c[3];
    cross(qv, v, t);
    for (int k = 0; k < 3; k++) t[k] *= 2.0;
    cross(qv, t, c);
    for (int k = 0; k < 3; k++) o[k] = t[k] * q[0] + c[k] + v[k];
}

// BoundaryType::world_aabb (:2122-2147): |R| applied to the half extents
// around the posed center
static void world_aabb(const Body &b, double mn[3], double mx[3]) {
    if (!b.dynamic) {
        for (int k = 0; k < 3; k++) {
            mn[k] = b.aabb_c[k] - b.aabb_h[k];
            mx[k] = b.aabb_c[k] + b.aabb_h[k];
        }
        return;
    }
    double R[9], hw[3], c[3];
    Rigid::quat_matrix(b.rb.q, R);
    for (double &r : R) r = std::fabs(r);
    Rigid::mv(R, b.aabb_h, hw);
    rotate(b.rb.q, b.aabb_c, c);
    for (int k = 0; k < 3; k++) c[k] = c[k] + b.rb.com[k];
    for (int k = 0; k < 3; k++) {
        mn[k] = c[k] - hw[k];
        mx[k] = c[k] + hw[k];
    }
}

// compute_diagonal (:884-946) from the device sums of one cell
static void diagonal(const Scene &S, const double *sums, double *r, double dt) {
    const Body &Ab = *S.bodies[(int)r[A]], &Bb = *S.bodies[(int)r[B]];
    r[VOL] = sums[0];
    for (int k = 0; k < 3; k++) {
        r[CEN + k] = sums[1 + k] / r[VOL];
        r[FA + k] = sums[4 + k];
        r[TA + k] = sums[7 + k];
        r[FB + k] = sums[10 + k];
        r[TB + k] = sums[13 + k];
    }
    double m[3], term_a = 0.0, term_b = 0.0;
    if (Ab.dynamic) {
        Rigid::mv(Ab.rb.Iinv_world, r + TA, m);
        term_a = dot(r + FA, r + FA) / Ab.rb.mass + dot(r + TA, m);
    }
    if (Bb.dynamic) {
        Rigid::mv(Bb.rb.Iinv_world, r + TB, m);
        term_b = dot(r + FB, r + FB) / Bb.rb.mass + dot(r + TB, m);
    }
    r[ACCP] = dt * dt * (term_a + term_b);
    // a_cc friction = 1/3 tr(dt T (K_a + K_b)), K = I/m - [c - com]x I^-1 [c - com]x
    const double fn = std::fmax(norm(r + FA), 1e-15);
    const double nrm[3] = {r[FA] / fn, r[FA + 1] / fn, r[FA + 2] / fn};
    double T[9];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) T[3 * i + j] = (i == j ? 1.0 : 0.0) - nrm[i] * nrm[j];
    double Ksum[9];
    const Body *pair[2] = {&Ab, &Bb};
    double Kb[2][9];
    for (int e = 0; e < 2; e++) {
        const Body &bb = *pair[e];
        if (!bb.dynamic) {
            for (double &v : Kb[e]) v = 0.0;
            continue;
        }
        double d[3], X[9], XI[9], XIX[9];
        sub(r + CEN, bb.rb.com, d);
        X[0] = 0.0; X[1] = -d[2]; X[2] = d[1];
        X[3] = d[2]; X[4] = 0.0; X[5] = -d[0];
        X[6] = -d[1]; X[7] = d[0]; X[8] = 0.0;
        Rigid::mm(X, bb.rb.Iinv_world, XI);
        Rigid::mm(XI, X, XIX);
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++)
                Kb[e][3 * i + j] = (i == j ? 1.0 : 0.0) / bb.rb.mass - XIX[3 * i + j];
    }
    for (int k = 0; k < 9; k++) Ksum[k] = Kb[0][k] + Kb[1][k];
    double dT[9], Pm[9];
    for (int k = 0; k < 9; k++) dT[k] = dt * T[k];
    Rigid::mm(dT, Ksum, Pm);
    r[ACCF] = 1. / 3. * (((0.0 + Pm[0]) + Pm[4]) + Pm[8]);
}

// solve_for_p_c (:186-637) over the active cells of one scene
struct Solver {
    const Scene &S;
    double *rec;
    size_t nc;
    int nb;
    double dt;
    std::vector<double> T, p, verr, lam, vtang, sp, sl, pj, lj, gp, gl, pp, vtn, dev, ef, et, vt,
        wt;
    std::vector<uint32_t> nact;

    Solver(const Scene &s, double *r, size_t n, double step)
        : S(s), rec(r), nc(n), nb((int)s.bodies.size()), dt(step), T(9 * n), p(n), verr(n),
          lam(3 * n), vtang(3 * n), sp(n), sl(3 * n), pj(n), lj(3 * n), gp(n), gl(3 * n), pp(n),
          vtn(n), dev(n), ef(3 * nb), et(3 * nb), vt(3 * nb), wt(3 * nb), nact(nb) {
        for (size_t c = 0; c < nc; c++) {  // tangential projection I - n n^T (:927-928)
            const double *r = cell(c);
            const double fn = std::fmax(norm(r + FA), 1e-15);
            const double nrm[3] = {r[FA] / fn, r[FA + 1] / fn, r[FA + 2] / fn};
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++)
                    T[9 * c + 3 * i + j] = (i == j ? 1.0 : 0.0) - nrm[i] * nrm[j];
        }
    }
    const double *cell(size_t c) const { return rec + c * kContactRec; }
    int body(size_t c, int side) const { return (int)cell(c)[side ? B : A]; }
    const Body &bd(int i) const { return *S.bodies[i]; }

    // accumulate_extra_forces (:948-986): each body over its cells in cell order
    void extra() {
        for (int bi = 0; bi < nb; bi++) {
            double f[3] = {0., 0., 0.}, t[3] = {0., 0., 0.};
            const bool dyn = bd(bi).dynamic;
            for (size_t c = 0; c < nc; c++) {
                const double *r = cell(c);
                const double *l = &lam[3 * c];
                for (int side = 0; side < 2; side++) {
                    if (body(c, side) != bi) continue;
                    const double *fs = r + (side ? FB : FA), *ts = r + (side ? TB : TA);
                    for (int k = 0; k < 3; k++)
                        f[k] += side ? p[c] * fs[k] - l[k] : p[c] * fs[k] + l[k];
                    if (dyn) {
                        double d[3], cr[3];
                        sub(r + CEN, bd(bi).rb.com, d);
                        cross(d, l, cr);
                        for (int k = 0; k < 3; k++)
                            t[k] += side ? p[c] * ts[k] - cr[k] : p[c] * ts[k] + cr[k];
                    }
                }
            }
            for (int k = 0; k < 3; k++) {
                ef[3 * bi + k] = f[k];
                et[3 * bi + k] = t[k];
            }
        }
    }
    // compute_trial_velocities (:988-1012); RigidBodyMotion::trial_velocity,
    // trial_angular_velocity (rigid_body_motion.rs:166-181)
    void trial() {
        for (int bi = 0; bi < nb; bi++) {
            const Body &b = bd(bi);
            if (!b.dynamic) {
                for (int k = 0; k < 3; k++) vt[3 * bi + k] = wt[3 * bi + k] = 0.0;
                continue;
            }
            double l[3];
            for (int k = 0; k < 3; k++) {
                vt[3 * bi + k] = b.rb.v[k] + dt / b.rb.mass * (b.rb.F[k] + ef[3 * bi + k]);
                l[k] = b.rb.L[k] + dt * (b.rb.T[k] + et[3 * bi + k]);
            }
            Rigid::mv(b.rb.Iinv_world, l, &wt[3 * bi]);
        }
    }
    // predicted_volume_error, predicted_tangential_velocity (:996-1033)
    void predict() {
        for (size_t c = 0; c < nc; c++) {
            const double *r = cell(c);
            const int a = body(c, 0), b = body(c, 1);
            const double *va = &vt[3 * a], *wa = &wt[3 * a], *vb = &vt[3 * b], *wb = &wt[3 * b];
            verr[c] = -r[VOL] +
                      dt * (((dot(va, r + FA) + dot(wa, r + TA)) + dot(vb, r + FB)) + dot(wb, r + TB));
            if (bd(a).dynamic && bd(b).dynamic) {
                double ra[3], rb[3], ca[3], cb[3], vr[3];
                sub(r + CEN, bd(a).rb.com, ra);
                sub(r + CEN, bd(b).rb.com, rb);
                cross(wa, ra, ca);
                cross(wb, rb, cb);
                for (int k = 0; k < 3; k++) vr[k] = ((va[k] + ca[k]) - vb[k]) - cb[k];
                Rigid::mv(&T[9 * c], vr, &vtang[3 * c]);
            } else {
                vtang[3 * c] = vtang[3 * c + 1] = vtang[3 * c + 2] = 0.0;
            }
        }
    }
    // projected_jacobi (:1035-1102)
    void jacobi() {
        for (int bi = 0; bi < nb; bi++) {
            uint32_t n = 0;
            for (size_t c = 0; c < nc; c++) {
                const int hits = (body(c, 0) == bi) + (body(c, 1) == bi);
                if (hits && (verr[c] < 0.0 || p[c] > 0.0)) n += (uint32_t)hits;
            }
            nact[bi] = n;
        }
        for (size_t c = 0; c < nc; c++) {
            const double *r = cell(c);
            const uint32_t nm = std::max(std::max(nact[body(c, 0)], nact[body(c, 1)]), 1u);
            const double n_max = (double)nm;
            const double alpha = kRelaxation / n_max / std::fmax(r[ACCP], kMinDiagonal);
            pj[c] = std::fmax(p[c] - alpha * verr[c], 0.0);
            const double lmax = kFriction * p[c] * norm(r + FA);
            const double d = std::fmax(r[ACCF], kMinDiagonal);
            double raw[3], star[3], nw[3];
            for (int k = 0; k < 3; k++) raw[k] = -vtang[3 * c + k] / d;
            const double rn = norm(raw);
            for (int k = 0; k < 3; k++) star[k] = rn > lmax ? lmax * raw[k] / rn : raw[k];
            for (int k = 0; k < 3; k++) nw[k] = lam[3 * c + k] + kRelaxation / n_max * star[k];
            const double nn = norm(nw);
            for (int k = 0; k < 3; k++) lj[3 * c + k] = nn > lmax ? lmax * nw[k] / nn : nw[k];
        }
    }
    // the merit inputs of the termination test (:310-334)
    void merit() {
        for (size_t c = 0; c < nc; c++) {
            const double *r = cell(c);
            pp[c] = p[c] * std::fmax(r[ACCP], kMinDiagonal);
            vtn[c] = norm(&vtang[3 * c]);
            dev[c] = std::fmax(r[ACCF], kMinDiagonal) *
                     (kFriction * p[c] * norm(r + FA) - norm(&lam[3 * c]));
        }
    }
    // mean_fischer_burmeister (utilities/mod.rs:27-41), summed in index order
    static double mean_fb(const std::vector<double> &a, const std::vector<double> &b) {
        double sum = 0.0;
        for (size_t i = 0; i < a.size(); i++)
            sum += std::fabs(a[i] + b[i] - std::sqrt(a[i] * a[i] + b[i] * b[i]));
        return sum / (double)a.size();
    }
    void refresh() {
        extra();
        trial();
        predict();
    }
    // the Jacobi step and its gradient (the negated step); returns |g_p|^2, |g_lambda|^2
    void step_gradient(double &g2p, double &g2l) {
        jacobi();
        for (size_t c = 0; c < nc; c++) {
            gp[c] = -(pj[c] - p[c]);
            for (int k = 0; k < 3; k++) gl[3 * c + k] = -(lj[3 * c + k] - lam[3 * c + k]);
        }
        g2p = 0.0;
        g2l = 0.0;
        for (size_t c = 0; c < nc; c++) g2p += gp[c] * gp[c];
        for (size_t c = 0; c < nc; c++) g2l += dot(&gl[3 * c], &gl[3 * c]);
    }

    uint32_t solve() {
        refresh();
        double prev_p2, prev_l2;
        step_gradient(prev_p2, prev_l2);
        for (size_t c = 0; c < nc; c++) {
            sp[c] = -gp[c];
            for (int k = 0; k < 3; k++) sl[3 * c + k] = -gl[3 * c + k];
        }
        p = pj;
        lam = lj;
        refresh();
        merit();
        uint32_t it = 0;
        while (!(mean_fb(pp, verr) < kFbPressure && mean_fb(vtn, dev) < kFbFriction)) {
            if (it >= kMaxIterations) break;
            double g2p, g2l;
            step_gradient(g2p, g2l);
            const double beta =
                (g2p + g2l) / std::fmax(prev_p2 + prev_l2, std::numeric_limits<double>::min());
            if (beta > 1.0) {  // restart from the Jacobi iterate
                std::fill(sp.begin(), sp.end(), 0.0);
                std::fill(sl.begin(), sl.end(), 0.0);
                p = pj;
                lam = lj;
            } else {
                // friction first: its cone bound takes the pressure before this update
                for (size_t c = 0; c < nc; c++) {
                    double ln[3], tp[3];
                    for (int k = 0; k < 3; k++) ln[k] = lj[3 * c + k] + beta * sl[3 * c + k];
                    Rigid::mv(&T[9 * c], ln, tp);
                    const double lmax = kFriction * p[c] * norm(cell(c) + FA);
                    const bool out = norm(tp) > lmax;
                    for (int k = 0; k < 3; k++) {
                        lam[3 * c + k] = out ? lj[3 * c + k] : tp[k];
                        sl[3 * c + k] = out ? 0.0 : beta * sl[3 * c + k] - gl[3 * c + k];
                    }
                }
                for (size_t c = 0; c < nc; c++) {
                    const double pn = pj[c] + beta * sp[c];
                    const bool neg = pn < 0.0;
                    p[c] = neg ? pj[c] : pn;
                    sp[c] = neg ? 0.0 : beta * sp[c] - gp[c];
                }
            }
            prev_p2 = g2p;
            prev_l2 = g2l;
            refresh();
            merit();
            it++;
        }
        extra();
        for (size_t c = 0; c < nc; c++) {
            rec[c * kContactRec + P] = p[c];
            for (int k = 0; k < 3; k++) rec[c * kContactRec + LAM + k] = lam[3 * c + k];
        }
        return it;
    }
};
}  // namespace contact

// contact handling of one scene before its bodies step
static void contact_step(Batch &B, Scene &S, double dt) {
    S.contact_rec.clear();
    S.contact_iters = 0;
    if (S.contact_degree <= 0 || S.bodies.size() < 2) return;
    bool dyn = false;
    for (auto &b : S.bodies) {
        if (b->kind != 1) return;
        dyn |= b->dynamic;
    }
    if (!dyn) return;
    // broad_phase (:70-92), generate_contact_cells (:98-115), subdivide_overlap
    // (:1813-1844)
    const double h = S.P.kernel_support_radius;
    const double target = h * contact::kCellSizeFactor;
    std::vector<ContactPairDev> pairs;
    uint64_t total = 0;
    const int nb = (int)S.bodies.size();
    for (int i = 0; i < nb; i++)
        for (int j = i + 1; j < nb; j++) {
            const Body &A = *S.bodies[i], &Bd = *S.bodies[j];
            if (!A.has_aabb || !Bd.has_aabb) continue;
            if (!A.dynamic && !Bd.dynamic) continue;
            double amn[3], amx[3], bmn[3], bmx[3];
            contact::world_aabb(A, amn, amx);
            contact::world_aabb(Bd, bmn, bmx);
            if (!(amn[0] <= bmx[0] && amx[0] >= bmn[0] && amn[1] <= bmx[1] && amx[1] >= bmn[1] &&
                  amn[2] <= bmx[2] && amx[2] >= bmn[2]))
                continue;
            ContactPairDev P{};
            P.a = i;
            P.b = j;
            double omx[3];
            for (int k = 0; k < 3; k++) {
                P.omn[k] = std::fmax(amn[k], bmn[k]);
                omx[k] = std::fmin(amx[k], bmx[k]);
            }
            if (!(P.omn[0] <= omx[0] && P.omn[1] <= omx[1] && P.omn[2] <= omx[2])) continue;
            uint64_t cells = 1;
            for (int k = 0; k < 3; k++) {
                const double ext = omx[k] - P.omn[k];
                const double c = std::ceil(ext / target);
                // f64 as usize saturates (NaN and negatives to 0), then max(1)
                const uint64_t ck = c > 0.0 ? (c < 4294967295.0 ? (uint64_t)c : 4294967295u) : 0;
                P.nn[k] = (uint32_t)std::max<uint64_t>(ck, 1);
                P.cs[k] = ext / (double)P.nn[k];
                cells *= P.nn[k];
            }
            if (total + cells > 0x7fffffffu)
                throw Error(SPH_ERR_UNSUPPORTED, "contact overlap exceeds 2^31 sub-cells");
            P.first = (uint32_t)total;
            total += cells;
            pairs.push_back(P);
        }
    if (pairs.empty()) return;
    std::vector<BodyDev> views;
    for (auto &b : S.bodies) views.push_back(b->view());
    S.cpairs.reserve(pairs.size());
    S.cbodies.reserve(views.size());
    S.cnact.reserve(1);
    B.put(S.cpairs.p, pairs.data(), pairs.size() * sizeof(ContactPairDev));
    B.put(S.cbodies.p, views.data(), views.size() * sizeof(BodyDev));
    ContactQuadDev Q{};
    Q.K = S.K;
    Q.degree = S.contact_degree;
    for (int k = 0; k < S.contact_degree; k++) {
        Q.x[k] = S.contact_x[k];
        Q.w[k] = S.contact_w[k];
    }
    // room for active cells: what the buffers hold (at least 4096), else all
    // of them after a first pass that only counts
    size_t cap = std::min<size_t>(total, std::max<size_t>(S.cidx.cap, 4096));
    size_t nc = 0;
    for (int pass = 0; pass < 2; pass++) {
        S.cidx.reserve(cap);
        S.csums.reserve(cap * kContactSums);
        LAUNCH(k_fill_u32, 1, B.s, (uint32_t *)S.cnact.p, (size_t)1, 0u);
        constexpr uint64_t kChunk = 1u << 20;  // workgroups per launch
        for (uint64_t base = 0; base < total; base += kChunk) {
            const uint64_t nblk = std::min<uint64_t>(kChunk, total - base);
            Prof::Rec r{SPH_PROF_CONTACT, (size_t)nblk, nullptr, nullptr};
            if (B.prof.on) B.prof.begin(B.s, r);
            hipLaunchKernelGGL(k_contact_cells, dim3((unsigned)nblk), dim3(kContactBlock), 0,
                               B.s, S.cpairs.p, (int)pairs.size(), S.cbodies.p, Q,
                               (uint32_t)base, S.cnact.p, (uint32_t)cap, S.cidx.p, S.csums.p);
            HIPCK(hipGetLastError());
            if (B.prof.on) B.prof.end(B.s, r);
        }
        unsigned int *hn = reinterpret_cast<unsigned int *>(B.pinned);
        HIPCK(hipMemcpyAsync(hn, S.cnact.p, sizeof(unsigned int), hipMemcpyDeviceToHost, B.s));
        B.sync();
        nc = *hn;
        if (nc > total) throw Error(SPH_ERR_HIP, "contact cell count out of range");
        if (nc <= cap) break;
        cap = nc;
    }
    if (nc > cap) throw Error(SPH_ERR_HIP, "contact cell count changed between passes");
    if (nc == 0) return;
    std::vector<uint32_t> idx(nc);
    std::vector<double> sums(nc * kContactSums);
    S.cidx.download(idx.data(), nc, B.s);
    S.csums.download(sums.data(), nc * kContactSums, B.s);
    B.sync();
    // the reference's cell order: candidate index (pair, then z, y, x)
    std::vector<uint32_t> order(nc);
    for (size_t k = 0; k < nc; k++) order[k] = (uint32_t)k;
    std::sort(order.begin(), order.end(), [&](uint32_t u, uint32_t v) { return idx[u] < idx[v]; });
    S.contact_rec.assign(nc * kContactRec, 0.0);
    for (size_t k = 0; k < nc; k++) {
        const uint32_t ci = idx[order[k]];
        size_t pi = 0;
        while (pi + 1 < pairs.size() && pairs[pi + 1].first <= ci) pi++;
        double *r = &S.contact_rec[k * kContactRec];
        r[contact::A] = (double)pairs[pi].a;
        r[contact::B] = (double)pairs[pi].b;
        contact::diagonal(S, &sums[(size_t)order[k] * kContactSums], r, dt);
    }
    contact::Solver solver(S, S.contact_rec.data(), nc, dt);
    S.contact_iters = solver.solve();
    for (int bi = 0; bi < nb; bi++) {  // add_force_and_torque (rigid_body_motion.rs:156-159)
        Body &b = *S.bodies[bi];
        if (!b.dynamic) continue;
        for (int k = 0; k < 3; k++) {
            b.rb.F[k] += solver.ef[3 * bi + k];
            b.rb.T[k] += solver.et[3 * bi + k];
        }
    }
}

// integrator + rigid bodies of every scene, then update (sph/mod.rs:125-144)
// a context steps alone: its failure is the call's error
static void throw_if_failed(const Batch &B) {
    if (const Scene *c = first_failed(B)) throw Error(c->failed, c->fail_msg);
}

static void step_batch(Batch &B) {
    layout(B);
    tune_begin_step(B, B.total_particles);
    BatchDev d = batch_dev(B);
    PLAUNCH(B, SPH_PROF_KERNELS, 0, kb_integrate_ec, B.total_pblocks, d);
    for (Scene *c : B.sc) {
        const double dt = c->dt;
        if (c->P.integrator == SPH_INTEGRATOR_TAKE_PREDICTED) {  // take_predicted.rs:14-27
            c->x.swap(c->ix);
            c->y.swap(c->iy);
            c->z.swap(c->iz);
            c->vx.swap(c->ivx);
            c->vy.swap(c->ivy);
            c->vz.swap(c->ivz);
        }
        contact_step(B, *c, dt);  // B::step_forward_in_time: contact, then the bodies
        for (auto &b : c->bodies)
            if (b->dynamic) {
                b->rb.step(dt);
                if (b->kind == 0) pose_samples(B, *c, *b);  // the samples follow the body
            }
        c->steps++;
        if (c->P.cfl)
            c->time += dt;
        else
            c->time = c->time_offset + (double)(c->steps - c->offset_steps) * c->P.time_increment;
    }
    update(B);
    tune_end_step(B, B.total_particles);
}

// SESPH and SESPHwSplitting: a state equation, no iterations
static bool state_equation(const sph_params &P) {
    return P.solver == SPH_SOLVER_SESPH || P.solver == SPH_SOLVER_SESPH_SPLITTING;
}

// a context's scene alone in the context's own batch
static Batch &solo(sph_ctx *c) {
    c->own.sc.assign(1, &c->scene);
    return c->own;
}

static void fill_measurement(sph_ctx *c, sph_measurement *o) {
    Scene &S = c->scene;
    Batch &B = solo(c);
    layout(B);
    std::vector<double> sums;
    measure_sums(B, sums);
    memset(o, 0, sizeof *o);
    const size_t n = S.n;
    o->time = S.time;
    o->time_step_size = S.dt;
    o->density = S.avg_density;
    o->density_error = n ? 100. * (sums[1] / (double)n - 1.) : 0.0;
    o->kinetic_energy = n ? sums[2] / (double)n : 0.0;
    o->solver_iterations = state_equation(S.P) ? 0 : S.last_iter;
    o->eqs1_iterations = S.P.solver == SPH_SOLVER_IISPHWOST ? S.eqs1_iter : 0;
    o->predicted_density_error = state_equation(S.P) ? 0.0 : S.last_err;
    o->steps = S.steps;
    o->num_particles = n;
}

static void validate_params(const sph_params &P) {
    if (!(P.kernel_support_radius > 0) || !(P.rest_density_grid_spacing > 0))
        throw Error(SPH_ERR_ARG, "kernel_support_radius and rest_density_grid_spacing must be > 0");
    if (P.solver < 0 || P.solver > 3) throw Error(SPH_ERR_ARG, "unknown solver");
    if (P.integrator < 0 || P.integrator > 3) throw Error(SPH_ERR_ARG, "unknown integrator");
    // PressureSolver::MANAGES_OWN_INTEGRATION pairing (setup/mod.rs:236-238)
    if ((P.solver == SPH_SOLVER_IISPHWOST) != (P.integrator == SPH_INTEGRATOR_TAKE_PREDICTED))
        throw Error(SPH_ERR_ARG, "IISPHwOST must be paired with TakePredicted and vice versa");
}

static void set_host_fluid(sph_ctx *c, size_t n, const double *xyz, const double *v,
                           const double *m, const uint32_t *fid) {
    Scene &S = c->scene;
    Batch &B = c->own;
    // neighbor gathers address a 32-byte record by a 32-bit byte offset
    if (n >= (1ull << 27))
        throw Error(SPH_ERR_UNSUPPORTED, "more than 2^27 particles in one system");
    reserve_scene(S, n);
    std::vector<double> X(n), Y(n), Z(n), U(n), W(n), T(n);
    for (size_t i = 0; i < n; i++) {
        X[i] = xyz[3 * i];
        Y[i] = xyz[3 * i + 1];
        Z[i] = xyz[3 * i + 2];
        U[i] = v ? v[3 * i] : 0.0;
        W[i] = v ? v[3 * i + 1] : 0.0;
        T[i] = v ? v[3 * i + 2] : 0.0;
    }
    S.x.upload(X.data(), n, B.s);
    S.y.upload(Y.data(), n, B.s);
    S.z.upload(Z.data(), n, B.s);
    S.vx.upload(U.data(), n, B.s);
    S.vy.upload(W.data(), n, B.s);
    S.vz.upload(T.data(), n, B.s);
    if (m) S.m.upload(m, n, B.s);
    else LAUNCH(k_fill_f64, n, B.s, S.m.p, n, 1.0);
    std::vector<uint32_t> ids(n), r(n);
    for (size_t i = 0; i < n; i++) {
        ids[i] = fid ? fid[i] : 0u;
        r[i] = (uint32_t)i;
    }
    S.fid.upload(ids.data(), n, B.s);
    S.ref.upload(r.data(), n, B.s);
    DBuf<double> *zero[] = {&S.ax, &S.ay, &S.az, &S.ix,  &S.iy, &S.iz,
                            &S.ivx, &S.ivy, &S.ivz, &S.p, &S.V};
    for (auto *q : zero) LAUNCH(k_fill_f64, n, B.s, q->p, n, 0.0);
    B.sync();
    S.n = n;
}

// download a sorted-order f64 field into caller order (by reference index)
static void download_field(sph_ctx *c, const DBuf<double> &f, std::vector<double> &out) {
    Scene &S = c->scene;
    const size_t n = S.n;
    std::vector<double> h(n);
    std::vector<uint32_t> r(n);
    f.download(h.data(), n, c->own.s);
    S.ref.download(r.data(), n, c->own.s);
    c->own.sync();
    out.assign(n, 0.0);
    for (size_t k = 0; k < n; k++) out[r[k]] = h[k];
}

// q4.w = V after a caller-provided volume replaced the computed one
static void pack_volume(Batch &B) {
    BatchDev d = batch_dev(B);
    PLAUNCH(B, SPH_PROF_KERNELS, 0, kb_pack_volume, B.total_pblocks, d);
}

static void check_solo(const sph_ctx *c) {
    if (c->scene.in_world) throw Error(SPH_ERR_STATE, "system is attached to a world");
}

// neighbors + boundary samples + pair cache + volume for a caller-provided
// state (per-trait entry points); the caller's index is the reference index
static void prepare_external(sph_ctx *c, size_t n, const double *x, const double *v,
                             const double *m) {
    check_solo(c);
    init_boundaries(c->own, c->scene);
    set_host_fluid(c, n, x, v, m, nullptr);
    Batch &B = solo(c);
    layout(B);
    prepare(B);
}

static void split3(const std::vector<double> &a, const std::vector<double> &b,
                   const std::vector<double> &cc, double *out) {
    for (size_t i = 0; i < a.size(); i++) {
        out[3 * i] = a[i];
        out[3 * i + 1] = b[i];
        out[3 * i + 2] = cc[i];
    }
}

// upload a caller-order f64 field into sorted order
static void upload_sorted(sph_ctx *c, const double *src, size_t stride, size_t comp,
                          DBuf<double> &dst) {
    Scene &S = c->scene;
    const size_t n = S.n;
    std::vector<uint32_t> r(n);
    S.ref.download(r.data(), n, c->own.s);
    c->own.sync();
    std::vector<double> h(n);
    for (size_t k = 0; k < n; k++) h[k] = src[stride * r[k] + comp];
    dst.upload(h.data(), n, c->own.s);
    c->own.sync();
}

}  // namespace sph

// ========================================================================
// C ABI
// ========================================================================
extern "C" {

const char *sph_last_error(void) { return sph::g_last_error.c_str(); }

int sph_ctx_create(const sph_params *params, sph_ctx **out) {
    return guarded([&] {
        if (!params || !out) throw Error(SPH_ERR_ARG, "null argument");
        validate_params(*params);
        check_device();
        std::unique_ptr<sph_ctx> c(new sph_ctx());
        Scene &S = c->scene;
        S.P = *params;
        const double h = params->kernel_support_radius, dxr = params->rest_density_grid_spacing;
        S.V0 = (dxr * dxr) * dxr;  // rest_volume = dx.powi(3) (sph/mod.rs:816)
        const double alpha = 16. / 3.14159265358979323846;
        S.K.h = h;
        S.K.wfac = alpha / ((h * h) * h);
        S.K.gfac = alpha / ((h * h) * (h * h));
        S.dthr = distance_threshold(h);
        S.reach = (int)std::ceil(h / h);
        S.dt = params->cfl ? 0.0 : params->time_increment;
        c->own.init();
        scratch_setup(c->own.device);
        reserve_scene(S, 1);
        *out = c.release();
    });
}

int sph_ctx_destroy(sph_ctx *c) {
    if (!c) return SPH_OK;
    return guarded([&] {
        if (c->scene.in_world) throw Error(SPH_ERR_STATE, "detach the system from its world first");
        delete c;
    });
}

int sph_ctx_clone(const sph_ctx *src, sph_ctx **out) {
    return guarded([&] {
        if (!src || !out) throw Error(SPH_ERR_ARG, "null argument");
        sph_ctx *c = nullptr;
        int st = sph_ctx_create(&src->scene.P, &c);
        if (st) throw Error(st, sph_last_error());
        std::unique_ptr<sph_ctx, int (*)(sph_ctx *)> guard(c, sph_ctx_destroy);
        const Scene &s = src->scene;
        Scene &d = c->scene;
        hipStream_t cs = c->own.s;
        HIPCK(hipStreamSynchronize(src->own.s));
        const size_t n = s.n;
        reserve_scene(d, n);
        DBuf<double> Scene::*f64s[] = {&Scene::x,  &Scene::y,   &Scene::z,   &Scene::vx,
                                       &Scene::vy, &Scene::vz,  &Scene::ax,  &Scene::ay,
                                       &Scene::az, &Scene::m,   &Scene::V,   &Scene::p,
                                       &Scene::ix, &Scene::iy,  &Scene::iz,  &Scene::ivx,
                                       &Scene::ivy, &Scene::ivz};
        for (auto f : f64s) (d.*f).copy_from(s.*f, n, cs);
        d.fid.copy_from(s.fid, n, cs);
        d.ref.copy_from(s.ref, n, cs);
        d.n = n;
        for (auto &bp : s.bodies) {
            std::unique_ptr<Body> b(new Body());
            b->kind = bp->kind;
            b->dynamic = bp->dynamic;
            b->rb = bp->rb;
            b->ns = bp->ns;
            b->spos = bp->spos;
            memcpy(b->applied, bp->applied, sizeof b->applied);
            b->has_aabb = bp->has_aabb;
            memcpy(b->aabb_c, bp->aabb_c, sizeof b->aabb_c);
            memcpy(b->aabb_h, bp->aabb_h, sizeof b->aabb_h);
            if (bp->kind == 1) {
                b->sdf.copy_from(bp->sdf, cs);
                b->vm.copy_from(bp->vm, cs);
            }
            if (bp->kind == 0 && bp->dynamic && bp->svol_set) {  // DynClone keeps the volumes
                b->svol_c.copy_from(bp->svol_c, bp->ns, cs);
                b->svol_set = true;
            }
            d.bodies.push_back(std::move(b));
        }
        c->own.sync();
        d.boundaries_initialized = false;
        init_boundaries(c->own, d);
        d.dt = s.dt;
        d.time = s.time;
        d.time_offset = s.time_offset;
        d.max_iter = s.max_iter;
        d.offset_steps = s.offset_steps;
        d.steps = s.steps;
        d.last_iter = s.last_iter;
        d.eqs1_iter = s.eqs1_iter;
        d.last_err = s.last_err;
        d.avg_density = s.avg_density;
        d.err_hist = s.err_hist;
        d.contact_degree = s.contact_degree;
        memcpy(d.contact_x, s.contact_x, sizeof d.contact_x);
        memcpy(d.contact_w, s.contact_w, sizeof d.contact_w);
        d.contact_rec = s.contact_rec;
        d.contact_iters = s.contact_iters;
        if (s.initialized) {  // rebuild neighbor/boundary state so the clone is self-contained
            Batch &B = solo(c);
            layout(B);
            prepare(B);
            B.sync();
        }
        d.initialized = s.initialized;
        *out = guard.release();
    });
}

int sph_neighbors_find(sph_ctx *c, double cell_size, double range, const double *pos, size_t n,
                       const double *samples, size_t m, uint64_t *offsets, uint64_t *indices,
                       size_t cap, size_t *total) {
    return guarded([&] {
        if (!c || (!pos && n) || (!samples && m) || !offsets || !total)
            throw Error(SPH_ERR_ARG, "null argument");
        if (!(cell_size > 0)) throw Error(SPH_ERR_ARG, "cell_size must be > 0");
        if (n >= 0xffffffffull || m >= 0xffffffffull)
            throw Error(SPH_ERR_UNSUPPORTED, "more than 2^32-1 points");
        check_solo(c);
        Batch &B = c->own;
        const int reach = (int)std::ceil(range / cell_size);
        if (reach < 0 || reach > 64) throw Error(SPH_ERR_UNSUPPORTED, "range / cell_size out of range");
        const double dthr = distance_threshold(range);
        auto soa = [&](const double *a, size_t cnt, DBuf<double> &X, DBuf<double> &Y,
                       DBuf<double> &Z) {
            std::vector<double> hx(cnt), hy(cnt), hz(cnt);
            for (size_t i = 0; i < cnt; i++) {
                hx[i] = a[3 * i];
                hy[i] = a[3 * i + 1];
                hz[i] = a[3 * i + 2];
            }
            X.upload(hx.data(), cnt, B.s);
            Y.upload(hy.data(), cnt, B.s);
            Z.upload(hz.data(), cnt, B.s);
            B.sync();
        };
        DBuf<double> qx, qy, qz, sx0, sy0, sz0, sx, sy, sz;
        soa(pos, n, qx, qy, qz);
        soa(samples, m, sx0, sy0, sz0);
        DBuf<int> sci, qci, bbs, bbq;
        sci.reserve(3 * std::max<size_t>(m, 1));
        qci.reserve(3 * std::max<size_t>(n, 1));
        bbs.reserve(6);
        bbq.reserve(6);
        cell_box(B, m, cell_size, sx0.p, sy0.p, sz0.p, sci.p, bbs.p);
        cell_box(B, n, cell_size, qx.p, qy.p, qz.p, qci.p, bbq.p);
        int b1[6], b2[6];
        HIPCK(hipMemcpyAsync(b1, bbs.p, sizeof b1, hipMemcpyDeviceToHost, B.s));
        HIPCK(hipMemcpyAsync(b2, bbq.p, sizeof b2, hipMemcpyDeviceToHost, B.s));
        B.sync();
        offsets[0] = 0;
        if (n == 0) {
            *total = 0;
            return;
        }
        if (m == 0) {
            for (size_t i = 0; i <= n; i++) offsets[i] = 0;
            *total = 0;
            return;
        }
        GridDev g = make_grid(b1, b1 + 3);
        DBuf<uint32_t> start, perm, part;
        cell_sort(B, m, sci.p, g, start, perm);
        g.start = start.p;
        g.mkey = perm.p;  // members in caller sample order (the bucket order)
        hash_partners(B, g, part, reach);
        sx.reserve(m);
        sy.reserve(m);
        sz.reserve(m);
        LAUNCH(k_gather<double>, m, B.s, m, perm.p, sx0.p, sx.p);
        LAUNCH(k_gather<double>, m, B.s, m, perm.p, sy0.p, sy.p);
        LAUNCH(k_gather<double>, m, B.s, m, perm.p, sz0.p, sz.p);
        DBuf<uint32_t> qcnt, off, idx;
        qcnt.reserve(n);
        off.reserve(n + 1);
        LAUNCH(k_neighbors<false>, n, B.s, n, qx.p, qy.p, qz.p, cell_size, reach, dthr, g, sx.p,
               sy.p, sz.p, (const uint32_t *)nullptr, qcnt.p, (uint32_t *)nullptr);
        uint32_t tot = scan_u32(B, qcnt.p, n, off.p);
        *total = tot;
        DBuf<uint64_t> off64;
        off64.reserve(n + 1);
        LAUNCH(k_widen_u32, n + 1, B.s, n + 1, off.p, off64.p);
        off64.download(offsets, n + 1, B.s);
        B.sync();
        if (tot > cap) throw Error(SPH_ERR_CAPACITY, "indices buffer too small");
        if (tot == 0) return;
        if (!indices) throw Error(SPH_ERR_ARG, "null indices");
        idx.reserve(tot);
        LAUNCH(k_neighbors<true>, n, B.s, n, qx.p, qy.p, qz.p, cell_size, reach, dthr, g, sx.p,
               sy.p, sz.p, off.p, (uint32_t *)nullptr, idx.p);
        DBuf<uint64_t> idx64;
        idx64.reserve(tot);
        LAUNCH(k_map_indices, tot, B.s, (size_t)tot, perm.p, idx.p, idx64.p);
        idx64.download(indices, tot, B.s);
        B.sync();
    });
}

int sph_neighbors_find_noalloc(sph_ctx *c, double cell_size, double range, const double *pos,
                               size_t n, const double *samples, size_t m, uint64_t *offsets,
                               uint64_t *indices, size_t cap, size_t *total) {
    return guarded([&] {
        if (!c || (!pos && n) || (!samples && m) || !offsets || !total)
            throw Error(SPH_ERR_ARG, "null argument");
        if (!(cell_size > 0)) throw Error(SPH_ERR_ARG, "cell_size must be > 0");
        if (n > kNaMaxPoints || m > kNaMaxPoints)
            throw Error(SPH_ERR_UNSUPPORTED,
                        "more than 262144 queries or samples (the static scratch limit)");
        check_solo(c);
        Batch &B = c->own;
        unsigned char *S = scratch_device(B.device);
        const int reach = (int)std::ceil(range / cell_size);
        if (reach < 0 || reach > 64) throw Error(SPH_ERR_UNSUPPORTED, "range / cell_size out of range");
        const double dthr = distance_threshold(range);
        int *bbox_s = reinterpret_cast<int *>(S + kNa.bbox_s);
        int *bbox_q = reinterpret_cast<int *>(S + kNa.bbox_q);
        unsigned int *flag = reinterpret_cast<unsigned int *>(S + kNa.flag);
        int *part = reinterpret_cast<int *>(S + kNa.part);
        int *sci = reinterpret_cast<int *>(S + kNa.sci);
        uint32_t *perm = reinterpret_cast<uint32_t *>(S + kNa.perm);
        double *ssx = reinterpret_cast<double *>(S + kNa.ssx);
        double *ssy = reinterpret_cast<double *>(S + kNa.ssy);
        double *ssz = reinterpret_cast<double *>(S + kNa.ssz);
        uint32_t *ucnt = reinterpret_cast<uint32_t *>(S + kNa.ucnt);
        uint32_t *ucur = reinterpret_cast<uint32_t *>(S + kNa.ucur);
        uint32_t *start = reinterpret_cast<uint32_t *>(S + kNa.start);
        uint32_t *sa = reinterpret_cast<uint32_t *>(S + kNa.scan_a);
        uint32_t *sb = reinterpret_cast<uint32_t *>(S + kNa.scan_b);
        uint32_t *sc = reinterpret_cast<uint32_t *>(S + kNa.scan_c);
        uint32_t *qcnt = reinterpret_cast<uint32_t *>(S + kNa.qcnt);
        uint32_t *off = reinterpret_cast<uint32_t *>(S + kNa.off);
        int *h = reinterpret_cast<int *>(B.pinned);  // 64 doubles of pinned host memory
        *total = 0;
        // occupied-cell boxes of the samples (with their cells) and of the queries
        auto box = [&](size_t cnt, const double *p, int *ci, int *bbox) {
            hipLaunchKernelGGL(k_init_bbox, dim3(1), dim3(64), 0, B.s, bbox);
            HIPCK(hipGetLastError());
            if (cnt == 0) return;
            LAUNCH(k_cells_aos, cnt, B.s, cnt, cell_size, p, ci, part);
            hipLaunchKernelGGL(k_fold_bbox, dim3(1), dim3(kBlock), 0, B.s, part,
                               (int)nblocks(cnt), bbox);
            HIPCK(hipGetLastError());
        };
        box(m, samples, sci, bbox_s);
        box(n, pos, nullptr, bbox_q);
        HIPCK(hipMemcpyAsync(h, bbox_s, 64, hipMemcpyDeviceToHost, B.s));  // both boxes
        B.sync();
        int b1[6], b2[6];
        std::memcpy(b1, h, sizeof b1);
        std::memcpy(b2, h + 8, sizeof b2);  // bbox_q sits 32 bytes after bbox_s
        if (n == 0 || m == 0) {  // offsets all zero
            LAUNCH(k_fill_u32, n + 1, B.s, off, n + 1, 0u);
            LAUNCH(k_widen_u32, n + 1, B.s, n + 1, off, offsets);
            B.sync();
            return;
        }
        // refuse boxes where two cells of one stencil share a hash bucket
        int lo[3], hi[3];
        for (int k = 0; k < 3; k++) {
            lo[k] = std::min(b1[k], b2[k]) - reach - 4;
            hi[k] = std::max(b1[3 + k], b2[3 + k]) + reach + 4;
        }
        GridDev hg = make_grid(lo, hi);
        LAUNCH(k_fill_u32, 1, B.s, (uint32_t *)flag, (size_t)1, 0u);
        LAUNCH(k_hash_check, (size_t)hg.dim[0] * hg.dim[1] * hg.dim[2], B.s, hg, 2 * reach, flag);
        // counting sort of the samples by cell, members in index order
        GridDev g = make_grid(b1, b1 + 3);
        const size_t ncell = (size_t)g.dim[0] * g.dim[1] * g.dim[2];
        if (ncell > kNaMaxCells)
            throw Error(SPH_ERR_UNSUPPORTED,
                        "occupied cell box of the samples exceeds 2^20 cells (the static "
                        "scratch limit)");
        LAUNCH(k_fill_u32, ncell, B.s, ucnt, ncell, 0u);
        LAUNCH(k_fill_u32, ncell, B.s, ucur, ncell, 0u);
        LAUNCH(k_cell_count, m, B.s, m, sci, g, ucnt);
        scan_launch(B.s, ucnt, ncell, start, sa, sb, sc);
        hipLaunchKernelGGL(k_scan_total, dim3(1), dim3(64), 0, B.s, ucnt, ncell, start);
        HIPCK(hipGetLastError());
        LAUNCH(k_cell_scatter, m, B.s, m, sci, g, start, ucur, perm);
        LAUNCH(k_cell_order, ncell, B.s, ncell, start, perm, (const uint32_t *)nullptr);
        g.start = start;
        LAUNCH(k_gather_aos, m, B.s, m, perm, samples, ssx, ssy, ssz);
        // counts, offsets, then the indices straight into the caller's buffer
        LAUNCH(k_neighbors_aos<false>, n, B.s, n, pos, cell_size, reach, dthr, g, ssx, ssy, ssz,
               (const uint32_t *)nullptr, qcnt, (const uint32_t *)nullptr, (uint64_t *)nullptr);
        scan_launch(B.s, qcnt, n, off, sa, sb, sc);
        uint32_t *last = reinterpret_cast<uint32_t *>(B.pinned);
        HIPCK(hipMemcpyAsync(&last[0], flag, sizeof(uint32_t), hipMemcpyDeviceToHost, B.s));
        HIPCK(hipMemcpyAsync(&last[1], off + n - 1, sizeof(uint32_t), hipMemcpyDeviceToHost, B.s));
        HIPCK(hipMemcpyAsync(&last[2], qcnt + n - 1, sizeof(uint32_t), hipMemcpyDeviceToHost, B.s));
        B.sync();
        if (last[0])
            throw Error(SPH_ERR_UNSUPPORTED,
                        "two cells of one neighbor stencil share a spatial-hash bucket; the "
                        "reference would report duplicate neighbors there");
        const uint64_t tot = (uint64_t)last[1] + last[2];
        if (tot > 0xffffffffull) throw Error(SPH_ERR_UNSUPPORTED, "more than 2^32 entries");
        hipLaunchKernelGGL(k_scan_total, dim3(1), dim3(64), 0, B.s, qcnt, n, off);
        HIPCK(hipGetLastError());
        LAUNCH(k_widen_u32, n + 1, B.s, n + 1, off, offsets);
        *total = (size_t)tot;
        if (tot > cap) {
            B.sync();
            throw Error(SPH_ERR_CAPACITY, "indices buffer too small");
        }
        if (tot > 0) {
            if (!indices) {
                B.sync();
                throw Error(SPH_ERR_ARG, "null indices");
            }
            LAUNCH(k_neighbors_aos<true>, n, B.s, n, pos, cell_size, reach, dthr, g, ssx, ssy,
                   ssz, off, (uint32_t *)nullptr, perm, indices);
        }
        B.sync();
    });
}

int sph_boundary_add_samples(sph_ctx *c, const double *pos, size_t m, const sph_rigid_init *rigid) {
    int idx = -1;
    int st = guarded([&] {
        if (!c || (!pos && m)) throw Error(SPH_ERR_ARG, "null argument");
        check_solo(c);
        if (c->scene.bodies.size() >= (size_t)kMaxBodiesPerScene)
            throw Error(SPH_ERR_UNSUPPORTED, "too many boundaries");
        if (rigid && c->scene.P.solver == SPH_SOLVER_IISPHWOST)  // SUPPORTS_DYNAMIC_BOUNDARIES (setup/mod.rs:242-244)
            throw Error(SPH_ERR_ARG, "IISPHwOST does not support dynamic boundaries");
        if (rigid && !(rigid->mass != 0.0)) throw Error(SPH_ERR_ARG, "rigid body mass cannot be zero");
        std::unique_ptr<Body> b(new Body());
        b->kind = 0;
        b->ns = m;
        b->spos.assign(pos, pos + 3 * m);
        if (rigid) {  // positions in the body frame (static_sample_boundary.rs:79-130)
            b->dynamic = true;
            b->rb.init(*rigid);
        }
        c->scene.bodies.push_back(std::move(b));
        c->scene.boundaries_initialized = false;
        idx = (int)c->scene.bodies.size() - 1;
    });
    return st ? st : idx;
}

int sph_boundary_add_volume_map(sph_ctx *c, const sph_lattice *sdf, const sph_lattice *vm,
                                const sph_rigid_init *rigid) {
    int idx = -1;
    int st = guarded([&] {
        if (!c || !sdf || !vm || !sdf->values || !sdf->present || !vm->values || !vm->present)
            throw Error(SPH_ERR_ARG, "null argument");
        check_solo(c);
        if (c->scene.bodies.size() >= (size_t)kMaxBodiesPerScene)
            throw Error(SPH_ERR_UNSUPPORTED, "too many boundaries");
        if (rigid && c->scene.P.solver == SPH_SOLVER_IISPHWOST)  // SUPPORTS_DYNAMIC_BOUNDARIES (setup/mod.rs:242-244)
            throw Error(SPH_ERR_ARG, "IISPHwOST does not support dynamic boundaries");
        if (rigid && !(rigid->mass != 0.0)) throw Error(SPH_ERR_ARG, "rigid body mass cannot be zero");
        std::unique_ptr<Body> b(new Body());
        b->kind = 1;
        b->sdf.load(*sdf, c->own.s);
        b->vm.load(*vm, c->own.s);
        if (rigid) {
            b->dynamic = true;
            b->rb.init(*rigid);
        }
        c->scene.bodies.push_back(std::move(b));
        idx = (int)c->scene.bodies.size() - 1;
    });
    return st ? st : idx;
}

int sph_boundary_set_aabb(sph_ctx *c, int bi, const double center[3],
                          const double half_extents[3]) {
    return guarded([&] {
        if (!c || !center || !half_extents) throw Error(SPH_ERR_ARG, "null argument");
        if (bi < 0 || (size_t)bi >= c->scene.bodies.size())
            throw Error(SPH_ERR_ARG, "bad boundary index");
        Body &b = *c->scene.bodies[bi];
        b.has_aabb = true;
        memcpy(b.aabb_c, center, sizeof b.aabb_c);
        memcpy(b.aabb_h, half_extents, sizeof b.aabb_h);
    });
}

int sph_contact_set_quadrature(sph_ctx *c, const double *gl_x, const double *gl_w, int degree) {
    return guarded([&] {
        if (!c || (degree > 0 && (!gl_x || !gl_w))) throw Error(SPH_ERR_ARG, "null argument");
        if (degree < 0 || degree > kContactMaxDegree)
            throw Error(SPH_ERR_UNSUPPORTED, "contact quadrature degree outside 0..16");
        Scene &S = c->scene;
        S.contact_degree = degree;
        for (int k = 0; k < kContactMaxDegree; k++) {
            S.contact_x[k] = k < degree ? gl_x[k] : 0.0;
            S.contact_w[k] = k < degree ? gl_w[k] : 0.0;
        }
    });
}

int sph_contact_stats(sph_ctx *c, uint32_t out[2]) {
    return guarded([&] {
        if (!c || !out) throw Error(SPH_ERR_ARG, "null argument");
        out[0] = (uint32_t)(c->scene.contact_rec.size() / kContactRec);
        out[1] = c->scene.contact_iters;
    });
}

int sph_contact_cells(sph_ctx *c, double *out, size_t cap, size_t *count) {
    return guarded([&] {
        if (!c || !count || (!out && cap)) throw Error(SPH_ERR_ARG, "null argument");
        const std::vector<double> &r = c->scene.contact_rec;
        const size_t n = r.size() / kContactRec;
        *count = n;
        if (n > cap) throw Error(SPH_ERR_CAPACITY, "buffer too small");
        if (n) memcpy(out, r.data(), r.size() * sizeof(double));
    });
}

int sph_fluid_upload(sph_ctx *c, size_t n, const double *x, const double *v, const double *m,
                     const uint32_t *fluid_id) {
    return guarded([&] {
        if (!c || (!x && n) || (!m && n)) throw Error(SPH_ERR_ARG, "null argument");
        check_solo(c);
        set_host_fluid(c, n, x, v, m, fluid_id);
        c->scene.initialized = false;
    });
}

int sph_system_init(sph_ctx *c) {
    return guarded([&] {
        if (!c) throw Error(SPH_ERR_ARG, "null argument");
        check_solo(c);
        Scene &S = c->scene;
        init_boundaries(c->own, S);
        S.dt = S.P.cfl ? 0.0 : S.P.time_increment;
        S.time = 0;
        S.steps = 0;
        S.time_offset = 0;
        S.offset_steps = 0;
        update(solo(c));
        throw_if_failed(c->own);
        S.initialized = true;
    });
}

int sph_step(sph_ctx *c, int nsteps, sph_measurement *out) {
    return guarded([&] {
        if (!c) throw Error(SPH_ERR_ARG, "null argument");
        check_solo(c);
        if (!c->scene.initialized) throw Error(SPH_ERR_STATE, "sph_system_init has not run");
        Batch &B = solo(c);
        for (int k = 0; k < nsteps; k++) {
            step_batch(B);
            throw_if_failed(B);
        }
        if (out) fill_measurement(c, out);
    });
}

int sph_fluid_len(const sph_ctx *c, size_t *n) {
    if (!c || !n) {
        set_error("null argument");
        return SPH_ERR_ARG;
    }
    *n = c->scene.n;
    return SPH_OK;
}

int sph_fluid_download(sph_ctx *c, double *x, double *v, double *a, double *V, double *p,
                       double *m, uint32_t *fluid_id) {
    return guarded([&] {
        if (!c) throw Error(SPH_ERR_ARG, "null argument");
        Scene &S = c->scene;
        std::vector<double> A, B_, C_;
        if (x) {
            download_field(c, S.x, A);
            download_field(c, S.y, B_);
            download_field(c, S.z, C_);
            split3(A, B_, C_, x);
        }
        if (v) {
            download_field(c, S.vx, A);
            download_field(c, S.vy, B_);
            download_field(c, S.vz, C_);
            split3(A, B_, C_, v);
        }
        if (a) {
            download_field(c, S.ax, A);
            download_field(c, S.ay, B_);
            download_field(c, S.az, C_);
            split3(A, B_, C_, a);
        }
        if (V) {
            download_field(c, S.V, A);
            memcpy(V, A.data(), A.size() * sizeof(double));
        }
        if (p) {
            download_field(c, S.p, A);
            memcpy(p, A.data(), A.size() * sizeof(double));
        }
        if (m) {
            download_field(c, S.m, A);
            memcpy(m, A.data(), A.size() * sizeof(double));
        }
        if (fluid_id) {
            const size_t n = S.n;
            std::vector<uint32_t> h(n), r(n);
            S.fid.download(h.data(), n, c->own.s);
            S.ref.download(r.data(), n, c->own.s);
            c->own.sync();
            for (size_t k = 0; k < n; k++) fluid_id[r[k]] = h[k];
        }
    });
}

int sph_measure(sph_ctx *c, sph_measurement *out) {
    return guarded([&] {
        if (!c || !out) throw Error(SPH_ERR_ARG, "null argument");
        fill_measurement(c, out);
    });
}

int sph_error_history(sph_ctx *c, double *out, size_t cap, size_t *count) {
    return guarded([&] {
        if (!c || !count) throw Error(SPH_ERR_ARG, "null argument");
        const auto &h = c->scene.err_hist;
        *count = h.size();
        if (out)
            for (size_t k = 0; k < std::min(cap, h.size()); k++) out[k] = h[k];
    });
}

int sph_body_state(sph_ctx *c, int bi, double out[22]) {
    return guarded([&] {
        if (!c || !out || bi < 0 || (size_t)bi >= c->scene.bodies.size())
            throw Error(SPH_ERR_ARG, "bad boundary index");
        const Rigid &r = c->scene.bodies[bi]->rb;
        memcpy(out, r.com, 3 * sizeof(double));
        memcpy(out + 3, r.q, 4 * sizeof(double));
        memcpy(out + 7, r.v, 3 * sizeof(double));
        memcpy(out + 10, r.L, 3 * sizeof(double));
        memcpy(out + 13, r.F, 3 * sizeof(double));
        memcpy(out + 16, r.T, 3 * sizeof(double));
        memcpy(out + 19, r.w, 3 * sizeof(double));
    });
}

// IntegrationScheme::POSITION_SLOTS / VELOCITY_SLOTS (integration_schemes/mod.rs:
// 23-42): TakePredicted one of each (take_predicted.rs:9-10), Verlet one
// position slot (verlet.rs:10-11), EulerCromer and ExplicitEuler none
static void slot_counts(const sph_params &P, int &npos, int &nvel) {
    npos = (P.integrator == SPH_INTEGRATOR_TAKE_PREDICTED || P.integrator == SPH_INTEGRATOR_VERLET) ? 1 : 0;
    nvel = P.integrator == SPH_INTEGRATOR_TAKE_PREDICTED ? 1 : 0;
}

int sph_fluid_slot_counts(sph_ctx *c, int *position_slots, int *velocity_slots) {
    return guarded([&] {
        if (!c || !position_slots || !velocity_slots) throw Error(SPH_ERR_ARG, "null argument");
        slot_counts(c->scene.P, *position_slots, *velocity_slots);
    });
}

int sph_fluid_upload_slots(sph_ctx *c, const double *xp, const double *vp) {
    return guarded([&] {
        if (!c) throw Error(SPH_ERR_ARG, "null argument");
        check_solo(c);
        Scene &S = c->scene;
        int npos, nvel;
        slot_counts(S.P, npos, nvel);
        const size_t n = S.n;
        Batch &B = c->own;
        // host arrays in the reference's order; the device keeps the particles
        // sorted once initialized, so position i takes reference index ref[i]
        auto put = [&](const double *h, DBuf<double> &X, DBuf<double> &Y, DBuf<double> &Z) {
            std::vector<double> a(n), b(n), d(n);
            for (size_t i = 0; i < n; i++) {
                a[i] = h[3 * i];
                b[i] = h[3 * i + 1];
                d[i] = h[3 * i + 2];
            }
            DBuf<double> *dst[3] = {&X, &Y, &Z};
            const std::vector<double> *src[3] = {&a, &b, &d};
            for (int k = 0; k < 3; k++) {
                DBuf<double> tmp;
                tmp.upload(src[k]->data(), n, B.s);
                LAUNCH(k_gather<double>, n, B.s, n, S.ref.p, tmp.p, dst[k]->p);
                B.sync();
            }
        };
        if (xp && npos) put(xp, S.ix, S.iy, S.iz);
        if (vp && nvel) put(vp, S.ivx, S.ivy, S.ivz);
    });
}

int sph_fluid_download_slots(sph_ctx *c, int *nslots, double *xp, double *vp) {
    return guarded([&] {
        if (!c || !nslots) throw Error(SPH_ERR_ARG, "null argument");
        Scene &S = c->scene;
        int nvel;
        slot_counts(S.P, *nslots, nvel);
        if (*nslots == 0) return;
        std::vector<double> A, B_, C_;
        if (xp) {
            download_field(c, S.ix, A);
            download_field(c, S.iy, B_);
            download_field(c, S.iz, C_);
            split3(A, B_, C_, xp);
        }
        if (vp && nvel) {
            download_field(c, S.ivx, A);
            download_field(c, S.ivy, B_);
            download_field(c, S.ivz, C_);
            split3(A, B_, C_, vp);
        }
    });
}

int sph_system_restore(sph_ctx *c, uint64_t time_steps_propagated, double current_time,
                       size_t nbodies, const uint8_t *present, const double *bodies) {
    return guarded([&] {
        if (!c || (nbodies && (!present || !bodies))) throw Error(SPH_ERR_ARG, "null argument");
        check_solo(c);
        Scene &S = c->scene;
        // the system as System::new leaves it (sample pseudo volumes at the
        // samples' initial pose), then the restored bodies
        init_boundaries(c->own, S);
        // BoundaryHandling::restore_from_checkpoint skips every boundary when
        // the entry count differs (volume_map_boundary.rs:1504-1518), and a
        // boundary whose entry does not match its kind (:2099-2120)
        if (nbodies == S.bodies.size()) {
            for (size_t b = 0; b < nbodies; b++) {
                Body &body = *S.bodies[b];
                if (!body.dynamic || !present[b]) continue;
                const double *s = bodies + 19 * b;
                Rigid &r = body.rb;
                memcpy(r.com, s, 3 * sizeof(double));
                // the serialized quaternion passes UnitQuaternion::from_quaternion
                // (rigid_body_motion.rs:273-289), which divides each coordinate
                // by the root of the squares summed in storage order (i, j, k, w)
                const double w = s[3], qi = s[4], qj = s[5], qk = s[6];
                const double nrm = sqrt(((qi * qi + qj * qj) + qk * qk) + w * w);
                r.q[0] = w / nrm;
                r.q[1] = qi / nrm;
                r.q[2] = qj / nrm;
                r.q[3] = qk / nrm;
                memcpy(r.v, s + 7, 3 * sizeof(double));
                memcpy(r.L, s + 10, 3 * sizeof(double));
                memcpy(r.F, s + 13, 3 * sizeof(double));
                memcpy(r.T, s + 16, 3 * sizeof(double));
                r.derive();  // RigidBodyMotion::restore_from_checkpoint (rigid_body_motion.rs:214-222)
                // samples on the body follow its restored pose (static_sample_boundary.rs:560-580)
                if (body.kind == 0) pose_samples(c->own, S, body);
            }
        }
        // both resume paths re-size the fluid's slots (sph/mod.rs:368-373,
        // 404-409; Fluid::resize_slots, fluid.rs:112-129): the integrator's
        // slots start at the origin and at zero again, and a checkpoint's
        // saved slots are not read
        DBuf<double> *slots[] = {&S.ix, &S.iy, &S.iz, &S.ivx, &S.ivy, &S.ivz};
        for (auto *q : slots) LAUNCH(k_fill_f64, S.n, c->own.s, q->p, S.n, 0.0);
        // continue_from_checkpoint (sph/mod.rs:362-378): counters, then update()
        S.dt = S.P.cfl ? 0.0 : S.P.time_increment;
        S.time = current_time;
        S.steps = time_steps_propagated;
        // a system built from a checkpoint takes its time as the offset
        // (new_boxed, sph/mod.rs:408-417)
        S.time_offset = current_time;
        S.offset_steps = time_steps_propagated;
        update(solo(c));
        throw_if_failed(c->own);
        S.initialized = true;
    });
}

int sph_system_status(const sph_ctx *c, int *status) {
    return guarded([&] {
        if (!c || !status) throw Error(SPH_ERR_ARG, "null argument");
        *status = c->scene.failed;
        if (c->scene.failed) set_error(c->scene.fail_msg.c_str());
    });
}

int sph_body_forces(sph_ctx *c, int bi, double out[6]) {
    return guarded([&] {
        if (!c || !out || bi < 0 || (size_t)bi >= c->scene.bodies.size())
            throw Error(SPH_ERR_ARG, "bad boundary index");
        memcpy(out, c->scene.bodies[bi]->applied, 6 * sizeof(double));
    });
}

int sph_profile(sph_ctx *c, int enable) {
    return guarded([&] {
        if (!c) throw Error(SPH_ERR_ARG, "null argument");
        c->own.sync();
        c->own.prof.drain();
        c->own.prof.reset();
        c->own.prof.on = enable != 0;
    });
}

int sph_profile_read(sph_ctx *c, int kernel, uint64_t *launches, uint64_t *units, double *ms) {
    return guarded([&] {
        if (!c || kernel < 0 || kernel >= SPH_PROF_KERNELS) throw Error(SPH_ERR_ARG, "bad argument");
        c->own.sync();
        c->own.prof.drain();
        if (launches) *launches = c->own.prof.launches[kernel];
        if (units) *units = c->own.prof.units[kernel];
        if (ms) *ms = c->own.prof.ms[kernel];
    });
}

int sph_volume(sph_ctx *c, size_t n, const double *x, double *V_out) {
    return guarded([&] {
        if (!c || (!x && n) || (!V_out && n)) throw Error(SPH_ERR_ARG, "null argument");
        prepare_external(c, n, x, nullptr, nullptr);
        std::vector<double> A;
        download_field(c, c->scene.V, A);
        if (n) memcpy(V_out, A.data(), n * sizeof(double));
        c->scene.initialized = false;
    });
}

int sph_boundary_find_samples(sph_ctx *c, int bi, size_t n, const double *x,
                              double *samples_out, uint8_t *flags_out) {
    return guarded([&] {
        if (!c || (!x && n) || bi < 0 || (size_t)bi >= c->scene.bodies.size())
            throw Error(SPH_ERR_ARG, "bad argument");
        Body &b = *c->scene.bodies[bi];
        if (b.kind != 1) throw Error(SPH_ERR_ARG, "boundary is not a volume map");
        prepare_external(c, n, x, nullptr, nullptr);
        std::vector<double> f[7];
        const DBuf<double> *src[7] = {&b.bx, &b.by, &b.bz, &b.bvx, &b.bvy, &b.bvz, &b.bvol};
        for (int k = 0; k < 7; k++) download_field(c, *src[k], f[k]);
        std::vector<uint8_t> fl(n), flr(n);
        std::vector<uint32_t> r(n);
        b.flag.download(fl.data(), n, c->own.s);
        c->scene.ref.download(r.data(), n, c->own.s);
        c->own.sync();
        for (size_t k = 0; k < n; k++) flr[r[k]] = fl[k];
        for (size_t i = 0; i < n; i++) {
            if (flags_out) flags_out[i] = flr[i];
            if (samples_out)
                for (int k = 0; k < 7; k++) samples_out[7 * i + k] = flr[i] ? f[k][i] : 0.0;
        }
        c->scene.initialized = false;
    });
}

int sph_nonpressure_accel(sph_ctx *c, size_t n, const double *x, const double *v, const double *m,
                          const double *V, double *a_out, double *body_ft_out) {
    return guarded([&] {
        if (!c || (n && (!x || !v || !m || !V || !a_out))) throw Error(SPH_ERR_ARG, "null argument");
        prepare_external(c, n, x, v, m);
        Scene &S = c->scene;
        upload_sorted(c, V, 1, 0, S.V);
        pack_volume(c->own);
        for (auto &b : S.bodies)
            for (int k = 0; k < 6; k++) b->applied[k] = 0;
        std::vector<Rigid> saved;
        for (auto &b : S.bodies) saved.push_back(b->rb);
        nonpressure(c->own);
        std::vector<double> A, B_, C_;
        download_field(c, S.ax, A);
        download_field(c, S.ay, B_);
        download_field(c, S.az, C_);
        split3(A, B_, C_, a_out);
        for (size_t bi = 0; bi < S.bodies.size(); bi++) {
            if (body_ft_out) memcpy(body_ft_out + 6 * bi, S.bodies[bi]->applied, 6 * sizeof(double));
            S.bodies[bi]->rb = saved[bi];  // stateless with respect to the bodies
        }
        S.initialized = false;
    });
}

int sph_pressure_solve(sph_ctx *c, size_t n, const double *x, const double *v, const double *m,
                       const double *V, double dt, double *a_inout, double *p_out,
                       double *body_ft_out, double *xpred_out, double *vpred_out,
                       sph_measurement *info_out) {
    return guarded([&] {
        if (!c || (n && (!x || !v || !m || !V || !a_inout))) throw Error(SPH_ERR_ARG, "null argument");
        prepare_external(c, n, x, v, m);
        Scene &S = c->scene;
        upload_sorted(c, V, 1, 0, S.V);
        pack_volume(c->own);
        upload_sorted(c, a_inout, 3, 0, S.ax);
        upload_sorted(c, a_inout, 3, 1, S.ay);
        upload_sorted(c, a_inout, 3, 2, S.az);
        for (auto &b : S.bodies)
            for (int k = 0; k < 6; k++) b->applied[k] = 0;
        std::vector<Rigid> saved;
        for (auto &b : S.bodies) saved.push_back(b->rb);
        const double keep_dt = S.dt;
        S.dt = dt;
        upload_views(c->own);
        S.failed = 0;
        S.fail_msg.clear();
        solve_pressure(c->own);
        S.dt = keep_dt;
        throw_if_failed(c->own);
        std::vector<double> A, B_, C_;
        download_field(c, S.ax, A);
        download_field(c, S.ay, B_);
        download_field(c, S.az, C_);
        split3(A, B_, C_, a_inout);
        if (p_out) {
            download_field(c, S.p, A);
            memcpy(p_out, A.data(), n * sizeof(double));
        }
        if (S.P.solver == SPH_SOLVER_IISPHWOST) {
            if (xpred_out) {
                download_field(c, S.ix, A);
                download_field(c, S.iy, B_);
                download_field(c, S.iz, C_);
                split3(A, B_, C_, xpred_out);
            }
            if (vpred_out) {
                download_field(c, S.ivx, A);
                download_field(c, S.ivy, B_);
                download_field(c, S.ivz, C_);
                split3(A, B_, C_, vpred_out);
            }
        }
        for (size_t bi = 0; bi < S.bodies.size(); bi++) {
            if (body_ft_out) memcpy(body_ft_out + 6 * bi, S.bodies[bi]->applied, 6 * sizeof(double));
            S.bodies[bi]->rb = saved[bi];
        }
        if (info_out) {
            memset(info_out, 0, sizeof *info_out);
            info_out->solver_iterations = state_equation(S.P) ? 0 : S.last_iter;
            info_out->eqs1_iterations = S.P.solver == SPH_SOLVER_IISPHWOST ? S.eqs1_iter : 0;
            info_out->predicted_density_error = state_equation(S.P) ? 0.0 : S.last_err;
            info_out->time_step_size = dt;
            info_out->num_particles = n;
        }
        S.initialized = false;
    });
}

// ---- worlds: many systems stepped together in shared launches ---------------
int sph_world_create(sph_world **out) {
    return guarded([&] {
        if (!out) throw Error(SPH_ERR_ARG, "null argument");
        check_device();
        std::unique_ptr<sph_world> w(new sph_world());
        w->batch.init();
        *out = w.release();
    });
}

int sph_world_destroy(sph_world *w) {
    if (!w) return SPH_OK;
    return guarded([&] {
        for (sph_ctx *c : w->members) c->scene.in_world = false;
        delete w;
    });
}

int sph_world_attach(sph_world *w, sph_ctx *c) {
    return guarded([&] {
        if (!w || !c) throw Error(SPH_ERR_ARG, "null argument");
        if (c->scene.in_world) throw Error(SPH_ERR_STATE, "system already attached to a world");
        if (!c->scene.initialized) throw Error(SPH_ERR_STATE, "sph_system_init has not run");
        if (!w->members.empty() && w->members[0]->scene.P.solver != c->scene.P.solver)
            throw Error(SPH_ERR_UNSUPPORTED, "all systems of a world use one pressure solver");
        c->own.release_scratch();  // the world's batch holds the launch-time buffers
        c->scene.in_world = true;
        w->members.push_back(c);
    });
}

int sph_world_detach(sph_world *w, sph_ctx *c) {
    return guarded([&] {
        if (!w || !c) throw Error(SPH_ERR_ARG, "null argument");
        auto it = std::find(w->members.begin(), w->members.end(), c);
        if (it == w->members.end()) throw Error(SPH_ERR_ARG, "system is not attached to this world");
        w->batch.sync();
        w->members.erase(it);
        c->scene.in_world = false;
    });
}

int sph_world_step(sph_world *w, int nsteps) {
    return guarded([&] {
        if (!w) throw Error(SPH_ERR_ARG, "null argument");
        Batch &B = w->batch;
        B.sc.clear();
        for (sph_ctx *c : w->members) B.sc.push_back(&c->scene);
        if (B.sc.empty()) return;
        for (int k = 0; k < nsteps; k++) {
            step_batch(B);
            // a failing member stops the call after this step: the others
            // have completed it and stay valid (sph_system_status names the
            // failed ones)
            for (size_t m = 0; m < w->members.size(); m++)
                if (w->members[m]->scene.failed) {
                    B.sync();
                    throw Error(w->members[m]->scene.failed,
                                "system " + std::to_string(m) + " of the world: " +
                                    w->members[m]->scene.fail_msg);
                }
        }
        B.sync();
    });
}

// ---- serving -------------------------------------------------------------------
static void server_stats(const sph_server *s, sph_server_stats *o) {
    o->queued = s->queued.size();
    o->admitted = (uint64_t)s->slots;
    o->resident = s->resident.size();
    o->finished = s->finished.size();
    o->failed = s->failed.size();
    o->bytes = s->used;
    o->world_steps = s->world_steps;
    o->device_bytes = g_device_bytes.load();
}

static void set_option(Batch &B, int option, int value);

// the server's world for one pressure solver, created on first use with the
// server's options
static sph_world *server_world(sph_server *s, int solver) {
    if (solver < 0 || solver >= kServerWorlds) throw Error(SPH_ERR_ARG, "unknown solver");
    sph_world *&w = s->worlds[solver];
    if (!w) {
        int e = sph_world_create(&w);
        if (e) throw Error(e, sph_last_error());
        for (const auto &o : s->options) set_option(w->batch, o.first, o.second);
    }
    return w;
}

static sph_world *world_of(sph_server *s, const sph_ctx *c) {
    return s->worlds[c->scene.P.solver];
}

int sph_server_create(uint64_t budget_bytes, int max_slots, sph_server **out) {
    return guarded([&] {
        if (!out || max_slots < 1) throw Error(SPH_ERR_ARG, "need an output and at least one slot");
        std::unique_ptr<sph_server> s(new sph_server());
        s->budget = budget_bytes;
        s->max_slots = max_slots;
        *out = s.release();
    });
}

int sph_server_destroy(sph_server *s) {
    if (!s) return SPH_OK;
    return guarded([&] {
        for (sph_world *w : s->worlds)
            if (w) (void)sph_world_destroy(w);
        delete s;
    });
}

int sph_server_submit(sph_server *s, uint64_t device_bytes, uint64_t steps, uint64_t *ticket) {
    return guarded([&] {
        if (!s || !ticket) throw Error(SPH_ERR_ARG, "null argument");
        if (steps == 0) throw Error(SPH_ERR_ARG, "a request needs at least one step");
        if (device_bytes > s->budget)
            throw Error(SPH_ERR_ARG, "request needs more device memory than the whole budget");
        s->queued.push_back({s->next_ticket, device_bytes, steps, 0, nullptr});
        *ticket = s->next_ticket++;
    });
}

int sph_server_admit(sph_server *s, uint64_t *ticket) {
    int got = 0;
    int st = guarded([&] {
        if (!s || !ticket) throw Error(SPH_ERR_ARG, "null argument");
        if (s->queued.empty()) return;
        const sph_server::Req &r = s->queued.front();
        if (s->slots >= s->max_slots || s->used + r.bytes > s->budget) return;
        s->admitted.push_back(r);
        s->queued.pop_front();
        s->slots++;
        s->used += r.bytes;
        *ticket = s->admitted.back().ticket;
        got = 1;
    });
    return st ? st : got;
}

int sph_server_attach(sph_server *s, uint64_t ticket, sph_ctx *ctx) {
    return guarded([&] {
        if (!s || !ctx) throw Error(SPH_ERR_ARG, "null argument");
        auto it = std::find_if(s->admitted.begin(), s->admitted.end(),
                               [&](const sph_server::Req &r) { return r.ticket == ticket; });
        if (it == s->admitted.end()) throw Error(SPH_ERR_STATE, "ticket is not awaiting a system");
        int e = sph_world_attach(server_world(s, ctx->scene.P.solver), ctx);
        if (e) throw Error(e, sph_last_error());
        sph_server::Req r = *it;
        r.ctx = ctx;
        s->admitted.erase(it);
        s->resident.push_back(r);
    });
}

// One step of every resident request: one world step per pressure solver that
// has resident requests. A request whose step fails (its pressure solve did
// not converge within its iteration cap) leaves its world for the failed
// queue and the others count the step as done; an error that is not confined
// to one request (a HIP error) fails every request of that world, the other
// worlds complete the step, and the error is returned.
int sph_server_step(sph_server *s, sph_server_stats *out) {
    return guarded([&] {
        if (!s) throw Error(SPH_ERR_ARG, "null argument");
        if (!s->resident.empty()) {
            int err = SPH_OK;
            std::string err_msg;
            for (sph_world *w : s->worlds) {
                if (!w || w->members.empty()) continue;
                int e = sph_world_step(w, 1);
                if (e == SPH_OK) continue;
                std::string msg = sph_last_error();
                bool confined = false;
                for (sph_ctx *c : w->members)
                    if (c->scene.failed) confined = true;
                if (confined) continue;
                // every request of this world fails with the world's error
                (void)hipStreamSynchronize(w->batch.s);
                std::vector<sph_server::Req> others;
                for (sph_server::Req &r : s->resident) {
                    if (world_of(s, r.ctx) != w) {
                        others.push_back(r);
                        continue;
                    }
                    r.ctx->scene.in_world = false;
                    s->failed.push_back({r, e});
                }
                w->members.clear();
                s->resident.swap(others);
                if (err == SPH_OK) {
                    err = e;
                    err_msg = msg;
                }
            }
            s->world_steps++;
            std::vector<sph_server::Req> keep;
            for (sph_server::Req &r : s->resident) {
                const int f = r.ctx->scene.failed;
                if (!f && ++r.done < r.steps) {
                    keep.push_back(r);
                    continue;
                }
                int e = sph_world_detach(world_of(s, r.ctx), r.ctx);
                if (e) throw Error(e, sph_last_error());
                if (f)
                    s->failed.push_back({r, f});
                else
                    s->finished.push_back(r);
            }
            s->resident.swap(keep);
            // a world's launch buffers only grow; once its resident particles
            // need less than a quarter of them (a large request left), free
            // them, and the next step reserves what it needs
            for (sph_world *w : s->worlds) {
                if (!w) continue;
                size_t n = 0;
                for (sph_ctx *c : w->members) n += c->scene.n;
                Batch &B = w->batch;
                if (B.perm.cap > 4 * std::max<size_t>(n, (size_t)1 << 16)) B.release_scratch();
            }
            if (err != SPH_OK) throw Error(err, err_msg);
        }
        if (out) server_stats(s, out);
    });
}

int sph_server_cancel(sph_server *s, uint64_t ticket) {
    return guarded([&] {
        if (!s) throw Error(SPH_ERR_ARG, "null argument");
        auto it = std::find_if(s->admitted.begin(), s->admitted.end(),
                               [&](const sph_server::Req &r) { return r.ticket == ticket; });
        if (it == s->admitted.end()) throw Error(SPH_ERR_STATE, "ticket is not awaiting a system");
        s->slots--;
        s->used -= it->bytes;
        s->admitted.erase(it);
    });
}

int sph_server_poll_failed(sph_server *s, uint64_t *ticket, sph_ctx **ctx, int *status) {
    int got = 0;
    int st = guarded([&] {
        if (!s || !ticket || !ctx || !status) throw Error(SPH_ERR_ARG, "null argument");
        if (s->failed.empty()) return;
        const sph_server::Failed f = s->failed.front();
        s->failed.pop_front();
        s->slots--;
        s->used -= f.req.bytes;
        *ticket = f.req.ticket;
        *ctx = f.req.ctx;
        *status = f.status;
        got = 1;
    });
    return st ? st : got;
}

int sph_server_poll(sph_server *s, uint64_t *ticket, sph_ctx **ctx) {
    int got = 0;
    int st = guarded([&] {
        if (!s || !ticket || !ctx) throw Error(SPH_ERR_ARG, "null argument");
        if (s->finished.empty()) return;
        const sph_server::Req r = s->finished.front();
        s->finished.pop_front();
        s->slots--;
        s->used -= r.bytes;
        *ticket = r.ticket;
        *ctx = r.ctx;
        got = 1;
    });
    return st ? st : got;
}

int sph_server_read_stats(const sph_server *s, sph_server_stats *out) {
    return guarded([&] {
        if (!s || !out) throw Error(SPH_ERR_ARG, "null argument");
        server_stats(s, out);
    });
}

int sph_world_size(const sph_world *w, size_t *systems, size_t *particles) {
    return guarded([&] {
        if (!w) throw Error(SPH_ERR_ARG, "null argument");
        size_t p = 0;
        for (sph_ctx *c : w->members) p += c->scene.n;
        if (systems) *systems = w->members.size();
        if (particles) *particles = p;
    });
}

int sph_world_profile(sph_world *w, int enable) {
    return guarded([&] {
        if (!w) throw Error(SPH_ERR_ARG, "null argument");
        w->batch.sync();
        w->batch.prof.drain();
        w->batch.prof.reset();
        w->batch.prof.on = enable != 0;
    });
}

int sph_world_profile_read(sph_world *w, int kernel, uint64_t *launches, uint64_t *units,
                           double *ms) {
    return guarded([&] {
        if (!w || kernel < 0 || kernel >= SPH_PROF_KERNELS) throw Error(SPH_ERR_ARG, "bad argument");
        w->batch.sync();
        w->batch.prof.drain();
        if (launches) *launches = w->batch.prof.launches[kernel];
        if (units) *units = w->batch.prof.units[kernel];
        if (ms) *ms = w->batch.prof.ms[kernel];
    });
}

// the errors set_option raises, without a batch
static void check_option(int option, int value) {
    const bool shape = option == SPH_OPT_SHAPE ||
                       (option >= SPH_OPT_SHAPE_VOLUME && option < SPH_OPT_SHAPE_VOLUME + kPairPasses);
    if (option != SPH_OPT_CACHE_GRADIENTS && option != SPH_OPT_AUTOTUNE && option != SPH_OPT_GRAPHS &&
        option != SPH_OPT_SWEEP && !shape)
        throw Error(SPH_ERR_ARG, "unknown option");
    if (option == SPH_OPT_AUTOTUNE && value < 0) throw Error(SPH_ERR_ARG, "steps per shape must be >= 0");
    if (option == SPH_OPT_SWEEP && value < 0) throw Error(SPH_ERR_ARG, "sweep option must be >= 0");
    if (shape && (value < 0 || value >= kPairVariants)) throw Error(SPH_ERR_ARG, "no such code shape");
}

static void set_option(Batch &B, int option, int value) {
    B.sync();
    if (option == SPH_OPT_CACHE_GRADIENTS) {
        B.grad_cache = value != 0;
        if (B.tuner.state == 2) B.tuner.state = 0;  // re-tune in the new mode
    } else if (option == SPH_OPT_AUTOTUNE) {
        if (value < 0) throw Error(SPH_ERR_ARG, "steps per shape must be >= 0");
        B.tuner.steps_per_shape = value;
        B.tuner.state = 0;
    } else if (option == SPH_OPT_GRAPHS) {
        B.use_graphs = value != 0;
        if (!B.use_graphs) B.drop_graphs();
    } else if (option == SPH_OPT_SWEEP) {
        if (value < 0) throw Error(SPH_ERR_ARG, "sweep option must be >= 0");
        B.sweep = value != 0;
        if (value > 1) B.ell_cap = (uint32_t)((value + 1) & ~1);  // rows come in pairs
    } else if (option == SPH_OPT_SHAPE) {
        if (value < 0 || value >= kPairVariants) throw Error(SPH_ERR_ARG, "no such code shape");
        for (int p = 0; p < kPairPasses; p++) B.variant[p] = value;
        B.tuner.steps_per_shape = 0;
        B.tuner.state = 0;
    } else if (option >= SPH_OPT_SHAPE_VOLUME && option < SPH_OPT_SHAPE_VOLUME + kPairPasses) {
        if (value < 0 || value >= kPairVariants) throw Error(SPH_ERR_ARG, "no such code shape");
        B.variant[option - SPH_OPT_SHAPE_VOLUME] = value;
        B.tuner.steps_per_shape = 0;
        B.tuner.state = 0;
    } else {
        throw Error(SPH_ERR_ARG, "unknown option");
    }
}

static int get_option(const Batch &B, int option) {
    if (option == SPH_OPT_CACHE_GRADIENTS) return B.grad_cache ? 1 : 0;
    if (option == SPH_OPT_AUTOTUNE) return B.tuner.steps_per_shape;
    if (option == SPH_OPT_TUNE_STATE) return B.tuner.state;
    if (option == SPH_OPT_GRAPHS) return B.use_graphs ? 1 : 0;
    if (option == SPH_OPT_GRAPH_COUNT) return (int)B.graphs.size();
    if (option == SPH_OPT_SWEEP) return B.sweep ? 1 : 0;
    if (option >= SPH_OPT_SHAPE_VOLUME && option < SPH_OPT_SHAPE_VOLUME + kPairPasses)
        return B.variant[option - SPH_OPT_SHAPE_VOLUME];
    throw Error(SPH_ERR_ARG, "unknown option");
}

int sph_get_option(const sph_ctx *c, int option, int *value) {
    return guarded([&] {
        if (!c || !value) throw Error(SPH_ERR_ARG, "null argument");
        if (option == SPH_OPT_MAX_ITERATIONS) {
            *value = (int)c->scene.max_iter;
            return;
        }
        *value = get_option(c->own, option);
    });
}

int sph_world_get_option(const sph_world *w, int option, int *value) {
    return guarded([&] {
        if (!w || !value) throw Error(SPH_ERR_ARG, "null argument");
        *value = get_option(w->batch, option);
    });
}

int sph_set_option(sph_ctx *c, int option, int value) {
    return guarded([&] {
        if (!c) throw Error(SPH_ERR_ARG, "null argument");
        if (option == SPH_OPT_MAX_ITERATIONS) {
            if (value < 1) throw Error(SPH_ERR_ARG, "the iteration cap must be >= 1");
            c->scene.max_iter = (uint32_t)value;
            return;
        }
        // every step and per-trait call rebuilds the pair cache before using it
        set_option(c->own, option, value);
    });
}

int sph_world_set_option(sph_world *w, int option, int value) {
    return guarded([&] {
        if (!w) throw Error(SPH_ERR_ARG, "null argument");
        set_option(w->batch, option, value);
    });
}

int sph_server_set_option(sph_server *s, int option, int value) {
    return guarded([&] {
        if (!s) throw Error(SPH_ERR_ARG, "null argument");
        check_option(option, value);
        for (sph_world *w : s->worlds)
            if (w) set_option(w->batch, option, value);
        s->options.emplace_back(option, value);
    });
}

int sph_quantity_at_positions(sph_ctx *c, int quantity, size_t nq, const float *pos,
                              float *out) {
    return guarded([&] {
        if (!c || (!pos && nq) || (!out && nq)) throw Error(SPH_ERR_ARG, "null argument");
        if (quantity < SPH_QUANTITY_SPEED || quantity > SPH_QUANTITY_KINETIC_ENERGY)
            throw Error(SPH_ERR_ARG, "unknown quantity");
        if (nq >= 0xffffffffull) throw Error(SPH_ERR_UNSUPPORTED, "more than 2^32-1 positions");
        check_solo(c);
        Scene &S = c->scene;
        if (!S.initialized) throw Error(SPH_ERR_STATE, "system not initialized");
        if (nq == 0) return;
        Batch &B = c->own;
        const size_t n = S.n;
        const double cs = S.P.kernel_support_radius;
        // f32 positions widened to f64 (sph/mod.rs:247-250)
        std::vector<double> hx(nq), hy(nq), hz(nq);
        for (size_t k = 0; k < nq; k++) {
            hx[k] = (double)pos[3 * k];
            hy[k] = (double)pos[3 * k + 1];
            hz[k] = (double)pos[3 * k + 2];
        }
        DBuf<double> qx, qy, qz;
        qx.upload(hx.data(), nq, B.s);
        qy.upload(hy.data(), nq, B.s);
        qz.upload(hz.data(), nq, B.s);
        DBuf<int> fci, bbf;
        fci.reserve(3 * std::max<size_t>(n, 1));
        bbf.reserve(6);
        cell_box(B, n, cs, S.x.p, S.y.p, S.z.p, fci.p, bbf.p);
        int b1[6];
        HIPCK(hipMemcpyAsync(b1, bbf.p, sizeof b1, hipMemcpyDeviceToHost, B.s));
        B.sync();
        // Duplicate neighbors in the reference come from two cells of one
        // stencil in one hash bucket; the static-sample grids carry their
        // bucket partners from set-up, the fluid's grid gets them here.
        std::vector<BodyDev> bv;
        std::vector<GridDev> sg;
        for (auto &b : S.bodies) {
            bv.push_back(b->view());
            sg.push_back(b->sg);
        }
        // the fluid's grid (one empty cell when there is no fluid)
        const int none[6] = {0, 0, 0, 0, 0, 0};
        GridDev g = make_grid(n ? b1 : none, n ? b1 + 3 : none + 3);
        DBuf<uint32_t> start, perm, part, mkey;
        cell_sort(B, n, fci.p, g, start, perm, S.ref.p);
        g.start = start.p;
        if (n) {
            hash_partners(B, g, part, S.reach);
            if (g.part) {  // merge keys: the reference index of each sorted member
                mkey.reserve(n);
                LAUNCH(k_gather<uint32_t>, n, B.s, n, perm.p, S.ref.p, mkey.p);
                g.mkey = mkey.p;
            }
        }
        DBuf<BodyDev> bodies;
        DBuf<GridDev> sgrid;
        if (!bv.empty()) {
            bodies.upload(bv.data(), bv.size(), B.s);
            sgrid.upload(sg.data(), sg.size(), B.s);
        }
        QueryConst C{};
        C.K = S.K;
        C.V0 = S.V0;
        C.cs = cs;
        C.dthr = S.dthr;
        C.dxr = S.P.rest_density_grid_spacing;
        C.vmax = 4.0 / 3.0 * 3.14159265358979323846 * ((cs * cs) * cs);  // ball_volume
        C.reach = S.reach;
        QueryFluid F{};
        F.x = S.x.p; F.y = S.y.p; F.z = S.z.p;
        F.vx = S.vx.p; F.vy = S.vy.p; F.vz = S.vz.p;
        F.m = S.m.p; F.V = S.V.p; F.p = S.p.p;
        F.perm = perm.p;
        F.g = g;
        DBuf<float> res;
        res.reserve(nq);
        const int nb = (int)bv.size();
        Prof::Rec r{SPH_PROF_QUANTITY, nq, nullptr, nullptr};
        if (B.prof.on) B.prof.begin(B.s, r);
#define QLAUNCH(Q) LAUNCH(k_quantity<Q>, nq, B.s, nq, qx.p, qy.p, qz.p, C, F, nb, bodies.p, sgrid.p, res.p)
        switch (quantity) {
            case SPH_QUANTITY_SPEED: QLAUNCH(kQSpeed); break;
            case SPH_QUANTITY_VOLUME: QLAUNCH(kQVolume); break;
            case SPH_QUANTITY_DENSITY: QLAUNCH(kQDensity); break;
            case SPH_QUANTITY_DENSITY_ERROR: QLAUNCH(kQDensityError); break;
            case SPH_QUANTITY_PRESSURE: QLAUNCH(kQPressure); break;
            default: QLAUNCH(kQKineticEnergy); break;
        }
#undef QLAUNCH
        if (B.prof.on) B.prof.end(B.s, r);
        res.download(out, nq, B.s);
        B.sync();
    });
}

int sph_volume_map_build(sph_ctx *c, const sph_lattice *sdf, double h, const double x_min[3],
                         const int64_t n[3], double dx, const double *gl_x, const double *gl_w,
                         int degree, double *values_out, uint8_t *present_out,
                         uint64_t *quadrature_nodes) {
    return guarded([&] {
        if (!sdf || !sdf->values || !sdf->present || !x_min || !n || !gl_x || !gl_w ||
            !values_out || !present_out)
            throw Error(SPH_ERR_ARG, "null argument");
        if (!(h > 0) || !(dx > 0)) throw Error(SPH_ERR_ARG, "h and dx must be > 0");
        if (degree < 1 || degree > kVmMaxDegree)
            throw Error(SPH_ERR_UNSUPPORTED, "quadrature degree outside 1..32");
        VmGrid g{};
        size_t nodes = 1;
        for (int d = 0; d < 3; d++) {
            if (n[d] <= 0 || n[d] > (1 << 20)) throw Error(SPH_ERR_ARG, "lattice n out of range");
            g.x_min[d] = x_min[d];
            g.n[d] = (int)n[d];
            nodes *= 3 * (size_t)n[d] + 1;
        }
        if (nodes >= ((size_t)1 << 31)) throw Error(SPH_ERR_UNSUPPORTED, "volume-map lattice too large");
        g.dx = dx;
        // a context lends its stream; without one the build runs on a stream of its own
        Batch local;
        if (!c) {
            check_device();
            local.init();
        }
        Batch &B = c ? c->own : local;
        LatticeHost L;
        L.load(*sdf, B.s);
        // node tables: radial, polar and azimuthal Gauss-Legendre points mapped
        // as ((b - a) x + (b + a)) / 2, with host sin/cos
        const double PI = 3.14159265358979323846;
        std::vector<double> r(degree), st(degree), ct(degree), sp(degree), cp(degree);
        for (int i = 0; i < degree; i++) {
            r[i] = 0.5 * ((h - 0.0) * gl_x[i] + (h + 0.0));
            const double th = 0.5 * ((PI - 0.0) * gl_x[i] + (PI + 0.0));
            const double ph = 0.5 * ((2. * PI - 0.0) * gl_x[i] + (2. * PI + 0.0));
            st[i] = std::sin(th);
            ct[i] = std::cos(th);
            sp[i] = std::sin(ph);
            cp[i] = std::cos(ph);
        }
        DBuf<double> tw, tr, tst, tct, tsp, tcp, values;
        tw.upload(gl_w, degree, B.s);
        tr.upload(r.data(), degree, B.s);
        tst.upload(st.data(), degree, B.s);
        tct.upload(ct.data(), degree, B.s);
        tsp.upload(sp.data(), degree, B.s);
        tcp.upload(cp.data(), degree, B.s);
        values.reserve(nodes);
        DBuf<uint8_t> status, present;
        status.reserve(nodes);
        present.reserve(nodes);
        DBuf<uint32_t> band, nband;
        band.reserve(nodes);
        nband.reserve(1);
        HIPCK(hipMemsetAsync(present.p, 0, nodes, B.s));
        LAUNCH(k_fill_u32, 1, B.s, nband.p, (size_t)1, 0u);
        const double bv = 4.0 / 3.0 * PI * ((h * h) * h);  // ball_volume (:34-36)
        LAUNCH(k_vm_classify, nodes, B.s, L.d, g, h, bv, values.p, status.p, band.p,
               (unsigned int *)nband.p);
        uint32_t nb = 0;
        nband.download(&nb, 1, B.s);
        B.sync();
        if (quadrature_nodes) *quadrature_nodes = nb;
        VmQuad Q{};
        Q.h = h;
        Q.bv = bv;
        Q.wfac = (16. / PI) / ((h * h) * h);
        Q.degree = degree;
        Q.w = tw.p;
        Q.r = tr.p;
        Q.st = tst.p;
        Q.ct = tct.p;
        Q.sp = tsp.p;
        Q.cp = tcp.p;
        // one workgroup per band node, in launches of at most 2^22 workgroups
        for (uint32_t off = 0; off < nb; off += (1u << 22)) {
            const uint32_t cnt = std::min<uint32_t>(nb - off, 1u << 22);
            Prof::Rec r{SPH_PROF_VM_QUADRATURE, cnt, nullptr, nullptr};
            if (B.prof.on) B.prof.begin(B.s, r);
            hipLaunchKernelGGL(k_vm_quadrature, dim3(cnt), dim3(kBlock), 0, B.s, L.d, g, Q,
                               band.p + off, values.p, status.p);
            HIPCK(hipGetLastError());
            if (B.prof.on) B.prof.end(B.s, r);
        }
        const size_t ncell = (size_t)n[0] * n[1] * n[2];
        LAUNCH(k_vm_cells, ncell, B.s, g, status.p, present.p);
        LAUNCH(k_vm_finish, nodes, B.s, nodes, present.p, values.p);
        values.download(values_out, nodes, B.s);
        present.download(present_out, nodes, B.s);
        B.sync();
    });
}

int sph_mesh_sdf_build(sph_ctx *c, const double *vertices, size_t nv, const uint32_t *triangles,
                       size_t nt, const double x_min[3], const int64_t n[3], double dx,
                       double prune, double *values_out, uint8_t *present_out) {
    return guarded([&] {
        if (!vertices || !triangles || !x_min || !n || !values_out || !present_out)
            throw Error(SPH_ERR_ARG, "null argument");
        if (nt == 0 || nt >= (1u << 31)) throw Error(SPH_ERR_ARG, "triangle count out of range");
        if (!(dx > 0) || !(prune >= 0)) throw Error(SPH_ERR_ARG, "dx must be > 0, prune >= 0");
        for (size_t i = 0; i < 3 * nt; i++)
            if (triangles[i] >= nv) throw Error(SPH_ERR_ARG, "triangle vertex index out of range");
        VmGrid g{};
        size_t nodes = 1;
        for (int d = 0; d < 3; d++) {
            if (n[d] <= 0 || n[d] > (1 << 20)) throw Error(SPH_ERR_ARG, "lattice n out of range");
            g.x_min[d] = x_min[d];
            g.n[d] = (int)n[d];
            nodes *= 3 * (size_t)n[d] + 1;
        }
        if (nodes >= ((size_t)1 << 31)) throw Error(SPH_ERR_UNSUPPORTED, "lattice too large");
        g.dx = dx;
        // pseudo-normals (Baerentzen and Aanaes 2005), host binary64: unit face
        // normals; vertices: incident faces in index order weighted by the
        // face angle at the vertex; edges: the sum of the incident faces'
        // normals over the unique vertex pairs; each normalized
        auto cross = [](const double *a, const double *b, double *o) {
            const double x = a[1] * b[2] - a[2] * b[1];
            const double y = a[2] * b[0] - a[0] * b[2];
            const double z = a[0] * b[1] - a[1] * b[0];
            o[0] = x;
            o[1] = y;
            o[2] = z;
        };
        auto dot = [](const double *a, const double *b) {
            return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2];
        };
        auto unit = [&](double *a) {
            const double l = std::sqrt(dot(a, a));
            if (l > 0.0)
                for (int k = 0; k < 3; k++) a[k] = a[k] / l;
        };
        std::vector<double> tri(9 * nt), fn(3 * nt, 0.0), vn(3 * nv, 0.0);
        for (size_t f = 0; f < nt; f++) {
            const double *a = vertices + 3 * triangles[3 * f], *b = vertices + 3 * triangles[3 * f + 1],
                         *cc = vertices + 3 * triangles[3 * f + 2];
            for (int k = 0; k < 3; k++) {
                tri[9 * f + k] = a[k];
                tri[9 * f + 3 + k] = b[k];
                tri[9 * f + 6 + k] = cc[k];
            }
            double e1[3], e2[3];
            for (int k = 0; k < 3; k++) {
                e1[k] = b[k] - a[k];
                e2[k] = cc[k] - a[k];
            }
            cross(e1, e2, &fn[3 * f]);
            unit(&fn[3 * f]);
        }
        for (size_t f = 0; f < nt; f++)
            for (int i = 0; i < 3; i++) {
                const double *p = vertices + 3 * triangles[3 * f + i],
                             *q = vertices + 3 * triangles[3 * f + (i + 1) % 3],
                             *r = vertices + 3 * triangles[3 * f + (i + 2) % 3];
                double e1[3], e2[3];
                for (int k = 0; k < 3; k++) {
                    e1[k] = q[k] - p[k];
                    e2[k] = r[k] - p[k];
                }
                const double l = std::sqrt(dot(e1, e1)) * std::sqrt(dot(e2, e2));
                double cs = l > 0.0 ? dot(e1, e2) / l : 1.0;
                cs = cs < -1.0 ? -1.0 : (cs > 1.0 ? 1.0 : cs);
                const double ang = std::acos(cs);
                double *o = &vn[3 * triangles[3 * f + i]];
                for (int k = 0; k < 3; k++) o[k] += ang * fn[3 * f + k];
            }
        for (size_t i = 0; i < nv; i++) unit(&vn[3 * i]);
        struct EdgeRef {
            uint32_t lo, hi, tri, local;
        };
        std::vector<EdgeRef> er(3 * nt);
        for (size_t f = 0; f < nt; f++)
            for (uint32_t i = 0; i < 3; i++) {
                const uint32_t a = triangles[3 * f + i], b = triangles[3 * f + (i + 1) % 3];
                er[3 * f + i] = {std::min(a, b), std::max(a, b), (uint32_t)f, i};
            }
        std::sort(er.begin(), er.end(), [](const EdgeRef &a, const EdgeRef &b) {
            if (a.lo != b.lo) return a.lo < b.lo;
            if (a.hi != b.hi) return a.hi < b.hi;
            if (a.tri != b.tri) return a.tri < b.tri;
            return a.local < b.local;
        });
        std::vector<uint32_t> te(3 * nt), tv(triangles, triangles + 3 * nt);
        std::vector<double> en(9 * nt, 0.0);
        size_t ne = 0;
        for (size_t k = 0; k < er.size(); k++) {
            if (k > 0 && (er[k].lo != er[k - 1].lo || er[k].hi != er[k - 1].hi)) ne++;
            te[3 * er[k].tri + er[k].local] = (uint32_t)ne;
            for (int d = 0; d < 3; d++) en[3 * ne + d] += fn[3 * er[k].tri + d];
        }
        ne++;
        for (size_t e = 0; e < ne; e++) unit(&en[3 * e]);
        Batch local;
        if (!c) {
            check_device();
            local.init();
        }
        Batch &B = c ? c->own : local;
        // spatially compact tiles: triangles ordered by the Morton code of their
        // centroid on a 1024^3 grid over the mesh bounds (ties by mesh index)
        double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
        double scale = 0.0;
        for (size_t i = 0; i < nv; i++)
            for (int d = 0; d < 3; d++) {
                lo[d] = std::fmin(lo[d], vertices[3 * i + d]);
                hi[d] = std::fmax(hi[d], vertices[3 * i + d]);
                scale = std::fmax(scale, std::fabs(vertices[3 * i + d]));
            }
        for (int d = 0; d < 3; d++) {
            scale = std::fmax(scale, std::fabs(x_min[d]));
            scale = std::fmax(scale, std::fabs(x_min[d] + (double)n[d] * dx));
        }
        auto spread = [](uint64_t x) {  // 10 bits -> every third bit
            x &= 0x3ff;
            x = (x | (x << 16)) & 0x30000ff;
            x = (x | (x << 8)) & 0x300f00f;
            x = (x | (x << 4)) & 0x30c30c3;
            x = (x | (x << 2)) & 0x9249249;
            return x;
        };
        std::vector<std::pair<uint64_t, uint32_t>> order(nt);
        for (size_t f = 0; f < nt; f++) {
            uint64_t code = 0;
            for (int d = 0; d < 3; d++) {
                const double cen = (tri[9 * f + d] + tri[9 * f + 3 + d] + tri[9 * f + 6 + d]) / 3.0;
                const double ext = hi[d] - lo[d];
                const double u = ext > 0.0 ? (cen - lo[d]) / ext : 0.0;
                const uint64_t q = (uint64_t)std::min(1023.0, std::max(0.0, u * 1024.0));
                code |= spread(q) << d;
            }
            order[f] = {code, (uint32_t)f};
        }
        std::sort(order.begin(), order.end());
        const size_t tiles = (nt + kMeshTile - 1) / kMeshTile;
        std::vector<double> stri(9 * nt), box(6 * tiles);
        std::vector<uint32_t> orig(nt);
        for (size_t i = 0; i < nt; i++) {
            orig[i] = order[i].second;
            for (int k = 0; k < 9; k++) stri[9 * i + k] = tri[9 * (size_t)order[i].second + k];
        }
        for (size_t tl = 0; tl < tiles; tl++) {
            double *bx = &box[6 * tl];
            for (int d = 0; d < 3; d++) {
                bx[d] = INFINITY;
                bx[3 + d] = -INFINITY;
            }
            for (size_t i = tl * kMeshTile; i < std::min(nt, (tl + 1) * kMeshTile); i++)
                for (int v = 0; v < 3; v++)
                    for (int d = 0; d < 3; d++) {
                        bx[d] = std::fmin(bx[d], stri[9 * i + 3 * v + d]);
                        bx[3 + d] = std::fmax(bx[3 + d], stri[9 * i + 3 * v + d]);
                    }
        }
        DBuf<double> dtri, dbox, dfn, dvn, den, values;
        DBuf<uint32_t> dtv, dte, dorig;
        DBuf<uint8_t> present;
        dtri.upload(stri.data(), stri.size(), B.s);
        dbox.upload(box.data(), box.size(), B.s);
        dorig.upload(orig.data(), orig.size(), B.s);
        dfn.upload(fn.data(), fn.size(), B.s);
        dvn.upload(vn.data(), vn.size(), B.s);
        den.upload(en.data(), 3 * ne, B.s);
        dtv.upload(tv.data(), tv.size(), B.s);
        dte.upload(te.data(), te.size(), B.s);
        values.reserve(nodes);
        present.reserve(nodes);
        HIPCK(hipMemsetAsync(present.p, 0, nodes, B.s));
        // distances are formed from coordinates of magnitude <= scale, so both the
        // tile-box and the triangle distance carry absolute errors far below
        // 1e-17 scale^2 (plus a relative 1e-6 in the test itself)
        MeshDev M{dtri.p, dorig.p, dbox.p, dtv.p, dte.p, dfn.p, dvn.p, den.p, (uint32_t)nt,
                  1e-17 * scale * scale};
        {
            Prof::Rec r{SPH_PROF_MESH_SDF, nodes, nullptr, nullptr};
            if (B.prof.on) B.prof.begin(B.s, r);
            LAUNCH(k_mesh_sdf, nodes, B.s, M, g, values.p);
            if (B.prof.on) B.prof.end(B.s, r);
        }
        const size_t ncell = (size_t)n[0] * n[1] * n[2];
        LAUNCH(k_sdf_cells, ncell, B.s, g, values.p, prune, present.p);
        LAUNCH(k_vm_finish, nodes, B.s, nodes, present.p, values.p);
        values.download(values_out, nodes, B.s);
        present.download(present_out, nodes, B.s);
        B.sync();
    });
}

int sph_device_bytes_held(uint64_t *bytes) {
    return guarded([&] {
        if (!bytes) throw Error(SPH_ERR_ARG, "null argument");
        *bytes = g_device_bytes.load();
    });
}

int sph_device_memory(size_t *free_bytes, size_t *total_bytes) {
    return guarded([&] {
        check_device();
        size_t f = 0, t = 0;
        HIPCK(hipMemGetInfo(&f, &t));
        if (free_bytes) *free_bytes = f;
        if (total_bytes) *total_bytes = t;
    });
}

}  // extern "C"
