// Host orchestration of the SPH hot path and its C ABI (include/sph_hip.h).
//
// A context owns one HIP stream and all device state of one system: fluid
// SoA arrays (binary64), the cell grid, the CSR neighbor list, per-body
// boundary data and reduction scratch. System::update (sph/mod.rs:674-731)
// is one sequence of launches on that stream; the host only reads back the
// few scalars the reference's control flow depends on (max speed for the CFL
// step, the occupied-cell box, the Jacobi error per iteration).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/sph_hip.h"
#include "sph_device.hpp"
#include "sph_kernels.hpp"

namespace sph {

// ------------------------------------------------------------------------
// errors
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
// optional per-kernel timing with HIP events on the context's stream
// ------------------------------------------------------------------------
struct Prof {
    bool on = false;
    struct Rec {
        int id;
        size_t units;
        hipEvent_t a, b;
    };
    std::vector<Rec> pending;
    std::vector<hipEvent_t> pool;
    uint64_t launches[SPH_PROF_KERNELS] = {0};
    uint64_t units[SPH_PROF_KERNELS] = {0};
    double ms[SPH_PROF_KERNELS] = {0};

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
    void drain() {  // caller synchronized the stream
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
        for (int k = 0; k < SPH_PROF_KERNELS; k++) {
            launches[k] = units[k] = 0;
            ms[k] = 0;
        }
    }
    ~Prof() {
        for (auto &r : pending) {
            (void)hipEventDestroy(r.a);
            (void)hipEventDestroy(r.b);
        }
        for (auto e : pool) (void)hipEventDestroy(e);
    }
};

// pair-cache arguments of the neighbor passes
#define ELL(c) (c)->nbr_cnt.p, (c)->slice_off.p, (c)->ell_j.p, (c)->ell_gx.p, (c)->ell_gy.p, (c)->ell_gz.p

// LAUNCH with event timing under profile id `pid` (units = particles)
#define PLAUNCH(ctx, pid, kernel, n, ...)                        \
    do {                                                         \
        Prof::Rec r_{pid, (size_t)(n), nullptr, nullptr};        \
        if ((ctx)->prof.on) (ctx)->prof.begin((ctx)->s, r_);     \
        LAUNCH(kernel, n, (ctx)->s, __VA_ARGS__);                \
        if ((ctx)->prof.on) (ctx)->prof.end((ctx)->s, r_);       \
    } while (0)

// ------------------------------------------------------------------------
// device buffer
// ------------------------------------------------------------------------
template <class T>
struct DBuf {
    T *p = nullptr;
    size_t cap = 0;
    DBuf() = default;
    DBuf(const DBuf &) = delete;
    DBuf &operator=(const DBuf &) = delete;
    ~DBuf() {
        if (p) (void)hipFree(p);
    }
    void reserve(size_t n) {
        if (n <= cap) return;
        size_t c = std::max<size_t>(n, cap + cap / 4);
        T *q = nullptr;
        HIPCK(hipMalloc((void **)&q, std::max<size_t>(c, 1) * sizeof(T)));
        if (p) (void)hipFree(p);
        p = q;
        cap = c;
    }
    // grow keeping the first `keep` elements
    void grow_keep(size_t n, size_t keep, hipStream_t s) {
        if (n <= cap) return;
        size_t c = std::max<size_t>(n, cap + cap / 4);
        T *q = nullptr;
        HIPCK(hipMalloc((void **)&q, std::max<size_t>(c, 1) * sizeof(T)));
        if (p && keep) HIPCK(hipMemcpyAsync(q, p, keep * sizeof(T), hipMemcpyDeviceToDevice, s));
        HIPCK(hipStreamSynchronize(s));
        if (p) (void)hipFree(p);
        p = q;
        cap = c;
    }
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
};

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
    void load(const sph_lattice &L, hipStream_t s) {
        for (int k = 0; k < 3; k++) {
            d.x_min[k] = L.x_min[k];
            d.x_max[k] = L.x_max[k];
            if (L.n[k] <= 0 || L.n[k] > (1 << 20)) throw Error(SPH_ERR_ARG, "lattice n out of range");
            d.n[k] = (int)L.n[k];
        }
        d.dx = L.dx;
        const size_t nx = 3 * (size_t)d.n[0] + 1, ny = 3 * (size_t)d.n[1] + 1,
                     nz = 3 * (size_t)d.n[2] + 1;
        const size_t nodes = nx * ny * nz;
        // a cell is evaluable iff all 32 of its nodes are present
        // (discretization.rs:455-463)
        static const int off[32][3] = {
            {0, 0, 0}, {3, 0, 0}, {0, 3, 0}, {3, 3, 0}, {0, 0, 3}, {3, 0, 3}, {0, 3, 3}, {3, 3, 3},
            {1, 0, 0}, {2, 0, 0}, {1, 3, 0}, {2, 3, 0}, {1, 0, 3}, {2, 0, 3}, {1, 3, 3}, {2, 3, 3},
            {0, 1, 0}, {0, 2, 0}, {3, 1, 0}, {3, 2, 0}, {0, 1, 3}, {0, 2, 3}, {3, 1, 3}, {3, 2, 3},
            {0, 0, 1}, {0, 0, 2}, {3, 0, 1}, {3, 0, 2}, {0, 3, 1}, {0, 3, 2}, {3, 3, 1}, {3, 3, 2}};
        const size_t ncell = (size_t)d.n[0] * d.n[1] * d.n[2];
        std::vector<uint8_t> ok(ncell, 1);
        std::vector<double> vals(nodes);
        for (size_t k = 0; k < nodes; k++) vals[k] = L.present[k] ? L.values[k] : 0.0;
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
        values.upload(vals.data(), nodes, s);
        cell_ok.upload(ok.data(), ncell, s);
        HIPCK(hipStreamSynchronize(s));
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
    std::vector<double> spos;            // caller order, [ns][3]
    DBuf<double> sx, sy, sz, svol;       // sorted by (cell, index)
    DBuf<uint32_t> sperm;                // sorted -> caller sample index
    DBuf<uint32_t> sgrid;                // cell start
    GridDev sg{};
    DBuf<uint32_t> soff, sidx, scnt;     // per fluid particle CSR
    // volume map
    LatticeHost sdf, vm;
    DBuf<uint8_t> flag;
    DBuf<double> bx, by, bz, bvx, bvy, bvz, bvol;
    double applied[6] = {0, 0, 0, 0, 0, 0};

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

}  // namespace sph

using namespace sph;

// ------------------------------------------------------------------------
// context
// ------------------------------------------------------------------------
struct sph_ctx {
    sph_params P{};
    double V0 = 0;
    KernelConst K{};
    double dthr = 0;
    int reach = 1;
    hipStream_t s = nullptr;
    int device = 0;

    size_t n = 0;
    // fluid state (sorted order)
    DBuf<double> x, y, z, vx, vy, vz, ax, ay, az, m, V, p;
    DBuf<uint32_t> fid, ref;
    // solver scratch
    DBuf<double> s_f, a_ff, pax, pay, paz, svx, svy, svz, sxp, syp, szp, ix, iy, iz, ivx, ivy, ivz;
    DBuf<double> tmp;   // gather scratch (f64)
    DBuf<uint32_t> tmpu;
    // grid + neighbors
    DBuf<int> ci;
    DBuf<uint32_t> perm, cell_cnt, cell_start, cursor;
    GridDev grid{};
    DBuf<uint32_t> nbr_off, nbr_idx, nbr_cnt;
    // wavefront-sliced ELL pair cache: neighbor index + grad W per slot
    DBuf<uint32_t> slice_len, slice_off, ell_j;
    DBuf<double> ell_gx, ell_gy, ell_gz;
    DBuf<uint32_t> scan_a, scan_b, scan_c;
    // reductions
    DBuf<double> partials, red;
    DBuf<int> bbox, ipart;
    DBuf<unsigned int> counters;
    DBuf<BodyDev> d_bodies;
    double *pinned = nullptr;  // small readbacks
    int checked_box[6] = {1, 1, 1, 0, 0, 0};
    bool hash_box_valid = false;

    std::vector<std::unique_ptr<Body>> bodies;
    bool boundaries_initialized = false;
    bool initialized = false;
    double dt = 0, time = 0;
    uint64_t steps = 0;
    uint32_t last_iter = 0, eqs1_iter = 0;
    double last_err = 0, avg_density = 0;
    std::vector<double> err_hist;
    Prof prof;
};

namespace sph {

static thread_local std::string g_last_error;

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

// exclusive scan of n uint32 into out (n+1 entries: out[n] = total); returns total
static uint32_t scan_u32(sph_ctx *c, const uint32_t *in, size_t n, uint32_t *out, bool read_total) {
    // out[0..n) = exclusive scan; total written at out[n]
    const size_t B = 1024;
    size_t nb1 = (n + B - 1) / B;
    if (n == 0) {
        uint32_t z = 0;
        HIPCK(hipMemcpyAsync(out, &z, sizeof z, hipMemcpyHostToDevice, c->s));
        HIPCK(hipStreamSynchronize(c->s));
        return 0;
    }
    c->scan_a.reserve(nb1 + 1);
    hipLaunchKernelGGL(k_scan_block, dim3((unsigned)nb1), dim3(512), 0, c->s, in, n, out,
                       c->scan_a.p);
    HIPCK(hipGetLastError());
    if (nb1 > 1) {
        size_t nb2 = (nb1 + B - 1) / B;
        c->scan_b.reserve(nb1 + 1);
        c->scan_c.reserve(nb2 + 1);
        hipLaunchKernelGGL(k_scan_block, dim3((unsigned)nb2), dim3(512), 0, c->s, c->scan_a.p,
                           nb1, c->scan_b.p, c->scan_c.p);
        HIPCK(hipGetLastError());
        if (nb2 > 1) {
            if (nb2 > B) throw Error(SPH_ERR_UNSUPPORTED, "scan longer than 2^30 elements");
            // third level: a single block
            DBuf<uint32_t> t3;
            t3.reserve(2);
            hipLaunchKernelGGL(k_scan_block, dim3(1), dim3(512), 0, c->s, c->scan_c.p, nb2,
                               c->scan_c.p, t3.p);
            HIPCK(hipGetLastError());
            hipLaunchKernelGGL(k_scan_add, dim3((unsigned)nb2), dim3(512), 0, c->s, c->scan_b.p,
                               nb1, c->scan_c.p);
            HIPCK(hipGetLastError());
            HIPCK(hipStreamSynchronize(c->s));
        }
        hipLaunchKernelGGL(k_scan_add, dim3((unsigned)nb1), dim3(512), 0, c->s, out, n,
                           c->scan_b.p);
        HIPCK(hipGetLastError());
    }
    // total = out[n-1] + in[n-1]
    uint32_t last[2];
    HIPCK(hipMemcpyAsync(&last[0], out + n - 1, sizeof(uint32_t), hipMemcpyDeviceToHost, c->s));
    HIPCK(hipMemcpyAsync(&last[1], in + n - 1, sizeof(uint32_t), hipMemcpyDeviceToHost, c->s));
    HIPCK(hipStreamSynchronize(c->s));
    uint64_t total = (uint64_t)last[0] + last[1];
    if (total > 0xffffffffull) throw Error(SPH_ERR_UNSUPPORTED, "more than 2^32 entries");
    uint32_t t32 = (uint32_t)total;
    HIPCK(hipMemcpyAsync(out + n, &t32, sizeof t32, hipMemcpyHostToDevice, c->s));
    HIPCK(hipStreamSynchronize(c->s));
    (void)read_total;
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

// refuse cell boxes where two cells that can share a stencil share a hash bucket
static void hash_guard(sph_ctx *c, const int lo[3], const int hi[3], int reach) {
    int elo[3], ehi[3];
    for (int k = 0; k < 3; k++) {
        elo[k] = lo[k] - reach;
        ehi[k] = hi[k] + reach;
    }
    if (c->hash_box_valid) {
        bool inside = true;
        for (int k = 0; k < 3; k++)
            inside = inside && elo[k] >= c->checked_box[k] && ehi[k] <= c->checked_box[3 + k];
        if (inside) return;
    }
    // check a margin around the box so small drifts do not re-trigger the check
    for (int k = 0; k < 3; k++) {
        elo[k] -= 4;
        ehi[k] += 4;
    }
    GridDev g = make_grid(elo, ehi);
    c->counters.reserve(4);
    LAUNCH(k_fill_u32, 4, c->s, (uint32_t *)c->counters.p, (size_t)4, 0u);
    size_t nc = (size_t)g.dim[0] * g.dim[1] * g.dim[2];
    LAUNCH(k_hash_check, nc, c->s, g, 2 * reach, c->counters.p);
    unsigned int flag = 0;
    HIPCK(hipMemcpyAsync(&flag, c->counters.p, sizeof flag, hipMemcpyDeviceToHost, c->s));
    HIPCK(hipStreamSynchronize(c->s));
    if (flag)
        throw Error(SPH_ERR_UNSUPPORTED,
                    "two cells of one neighbor stencil share a spatial-hash bucket; the "
                    "reference would report duplicate neighbors there");
    for (int k = 0; k < 3; k++) {
        c->checked_box[k] = elo[k];
        c->checked_box[3 + k] = ehi[k];
    }
    c->hash_box_valid = true;
}

// counting sort of `cnt` points (with cell coords ci) into (cell, key) order;
// fills perm (sorted -> current index) and cell_start over grid g
static void cell_sort(sph_ctx *c, size_t cnt, const int *ci, const GridDev &g,
                      DBuf<uint32_t> &cell_cnt, DBuf<uint32_t> &start, DBuf<uint32_t> &perm,
                      const uint32_t *order_key) {
    size_t ncell = (size_t)g.dim[0] * g.dim[1] * g.dim[2];
    cell_cnt.reserve(ncell);
    start.reserve(ncell + 1);
    c->cursor.reserve(ncell);
    perm.reserve(cnt);
    LAUNCH(k_fill_u32, ncell, c->s, cell_cnt.p, ncell, 0u);
    LAUNCH(k_fill_u32, ncell, c->s, c->cursor.p, ncell, 0u);
    LAUNCH(k_cell_count, cnt, c->s, cnt, ci, g, cell_cnt.p);
    scan_u32(c, cell_cnt.p, ncell, start.p, false);
    LAUNCH(k_cell_scatter, cnt, c->s, cnt, ci, g, start.p, c->cursor.p, perm.p);
    LAUNCH(k_cell_order, ncell, c->s, ncell, start.p, perm.p, order_key);
}

template <class T>
static void permute(sph_ctx *c, DBuf<T> &field, DBuf<T> &scratch, size_t cnt) {
    scratch.reserve(cnt);
    LAUNCH(k_gather<T>, cnt, c->s, cnt, c->perm.p, field.p, scratch.p);
    std::swap(field.p, scratch.p);
    std::swap(field.cap, scratch.cap);
}

static void reserve_fluid(sph_ctx *c, size_t n) {
    DBuf<double> *f64s[] = {&c->x,   &c->y,   &c->z,   &c->vx,  &c->vy,  &c->vz,  &c->ax,
                            &c->ay,  &c->az,  &c->m,   &c->V,   &c->p,   &c->s_f, &c->a_ff,
                            &c->pax, &c->pay, &c->paz, &c->svx, &c->svy, &c->svz, &c->sxp,
                            &c->syp, &c->szp, &c->ix,  &c->iy,  &c->iz,  &c->ivx, &c->ivy,
                            &c->ivz, &c->tmp};
    for (auto *b : f64s) b->reserve(n);
    c->fid.reserve(n);
    c->ref.reserve(n);
    c->tmpu.reserve(n);
    c->ci.reserve(3 * n);
    c->perm.reserve(n);
    c->nbr_off.reserve(n + 1);
    c->nbr_cnt.reserve(n);
    c->partials.reserve((size_t)(nblocks(n) + 1) * kMaxBodies * 6 + 64);
    c->red.reserve(kMaxBodies * 6 + 64);
}

// cell coordinates of `cnt` points and their occupied-cell box into bbox (device)
static void cell_box(sph_ctx *c, size_t cnt, double cs, const double *X, const double *Y,
                     const double *Z, int *ci, int *bbox) {
    hipLaunchKernelGGL(k_init_bbox, dim3(1), dim3(64), 0, c->s, bbox);
    HIPCK(hipGetLastError());
    if (cnt == 0) return;
    unsigned nbk = nblocks(cnt);
    c->ipart.reserve((size_t)nbk * 6);
    LAUNCH(k_cells, cnt, c->s, cnt, cs, X, Y, Z, ci, c->ipart.p);
    hipLaunchKernelGGL(k_fold_bbox, dim3(1), dim3(kBlock), 0, c->s, c->ipart.p, (int)nbk, bbox);
    HIPCK(hipGetLastError());
}

static void upload_bodies(sph_ctx *c) {
    std::vector<BodyDev> v;
    for (auto &b : c->bodies) v.push_back(b->view());
    c->d_bodies.reserve(std::max<size_t>(v.size(), 1));
    if (!v.empty())
        HIPCK(hipMemcpyAsync(c->d_bodies.p, v.data(), v.size() * sizeof(BodyDev),
                             hipMemcpyHostToDevice, c->s));
}

// sort the fluid by (cell, reference index), build the cell grid and the
// CSR neighbor list (System::update -> N::find_samples, sph/mod.rs:701-706)
static void fluid_neighbors(sph_ctx *c) {
    const size_t n = c->n;
    c->bbox.reserve(6);
    cell_box(c, n, c->P.kernel_support_radius, c->x.p, c->y.p, c->z.p, c->ci.p, c->bbox.p);
    int bb[6];
    HIPCK(hipMemcpyAsync(bb, c->bbox.p, sizeof bb, hipMemcpyDeviceToHost, c->s));
    HIPCK(hipStreamSynchronize(c->s));
    if (n == 0) {
        for (int k = 0; k < 3; k++) {
            bb[k] = 0;
            bb[3 + k] = 0;
        }
    }
    hash_guard(c, bb, bb + 3, c->reach);
    c->grid = make_grid(bb, bb + 3);
    cell_sort(c, n, c->ci.p, c->grid, c->cell_cnt, c->cell_start, c->perm, c->ref.p);
    c->grid.start = c->cell_start.p;
    permute(c, c->x, c->tmp, n);
    permute(c, c->y, c->tmp, n);
    permute(c, c->z, c->tmp, n);
    permute(c, c->vx, c->tmp, n);
    permute(c, c->vy, c->tmp, n);
    permute(c, c->vz, c->tmp, n);
    permute(c, c->m, c->tmp, n);
    permute(c, c->fid, c->tmpu, n);
    permute(c, c->ref, c->tmpu, n);
    // neighbor counts and per-slice maxima; the pair cache is filled after the
    // boundary samples exist (pairs_and_volume)
    const size_t nsl = (n + 63) / 64;
    c->slice_len.reserve(nsl + 1);
    c->slice_off.reserve(nsl + 1);
    PLAUNCH(c, SPH_PROF_NEIGHBOR_COUNT, k_count_ell, n, n, c->x.p, c->y.p, c->z.p,
            c->P.kernel_support_radius, c->reach, c->dthr, c->grid, c->nbr_cnt.p,
            c->slice_len.p);
    uint32_t total = scan_u32(c, c->slice_len.p, nsl, c->slice_off.p, true);
    uint64_t slots = (uint64_t)total * 64;
    if (slots > 0xffffffffull) throw Error(SPH_ERR_UNSUPPORTED, "neighbor cache exceeds 2^32 slots");
    c->ell_j.reserve(std::max<uint64_t>(slots, 1));
    c->ell_gx.reserve(std::max<uint64_t>(slots, 1));
    c->ell_gy.reserve(std::max<uint64_t>(slots, 1));
    c->ell_gz.reserve(std::max<uint64_t>(slots, 1));
}

// neighbor indices, then grad W into the pair cache and the volume (get_volume)
static void pairs_and_volume(sph_ctx *c) {
    const size_t n = c->n;
    PLAUNCH(c, SPH_PROF_PAIRS_VOLUME, k_fill_ell, n, n, c->P.kernel_support_radius, c->reach,
            c->dthr, c->grid, c->x.p, c->y.p, c->z.p, c->slice_off.p, c->ell_j.p);
    PLAUNCH(c, SPH_PROF_PAIRS_VOLUME, k_pairs_volume, n, n, c->K, c->V0, c->x.p, c->y.p, c->z.p,
            ELL(c), c->d_bodies.p, (int)c->bodies.size(), c->V.p);
}

// StaticSampleBoundary samples: sort once by (cell, index) and build their grid
static void prepare_samples(sph_ctx *c, Body &b) {
    size_t m = b.ns;
    DBuf<double> hx, hy, hz;
    std::vector<double> X(m), Y(m), Z(m);
    for (size_t k = 0; k < m; k++) {
        X[k] = b.spos[3 * k];
        Y[k] = b.spos[3 * k + 1];
        Z[k] = b.spos[3 * k + 2];
    }
    hx.upload(X.data(), m, c->s);
    hy.upload(Y.data(), m, c->s);
    hz.upload(Z.data(), m, c->s);
    DBuf<int> sci;
    sci.reserve(3 * std::max<size_t>(m, 1));
    c->bbox.reserve(6);
    cell_box(c, m, c->P.kernel_support_radius, hx.p, hy.p, hz.p, sci.p, c->bbox.p);
    int bb[6];
    HIPCK(hipMemcpyAsync(bb, c->bbox.p, sizeof bb, hipMemcpyDeviceToHost, c->s));
    HIPCK(hipStreamSynchronize(c->s));
    if (m == 0)
        for (int k = 0; k < 6; k++) bb[k] = 0;
    b.sg = make_grid(bb, bb + 3);
    DBuf<uint32_t> cnt;
    cell_sort(c, m, sci.p, b.sg, cnt, b.sgrid, b.sperm, nullptr);
    b.sg.start = b.sgrid.p;
    b.sx.reserve(m);
    b.sy.reserve(m);
    b.sz.reserve(m);
    b.svol.reserve(m);
    LAUNCH(k_gather<double>, m, c->s, m, b.sperm.p, hx.p, b.sx.p);
    LAUNCH(k_gather<double>, m, c->s, m, b.sperm.p, hy.p, b.sy.p);
    LAUNCH(k_gather<double>, m, c->s, m, b.sperm.p, hz.p, b.sz.p);
    // pseudo volumes (BoundaryType::initialize)
    DBuf<uint32_t> off, idx, scnt;
    off.reserve(m + 1);
    scnt.reserve(std::max<size_t>(m, 1));
    LAUNCH(k_neighbors<false>, m, c->s, m, b.sx.p, b.sy.p, b.sz.p, c->P.kernel_support_radius,
           c->reach, c->dthr, b.sg, b.sx.p, b.sy.p, b.sz.p, (const uint32_t *)nullptr, scnt.p,
           (uint32_t *)nullptr);
    uint32_t total = scan_u32(c, scnt.p, m, off.p, true);
    idx.reserve(std::max<uint32_t>(total, 1));
    LAUNCH(k_neighbors<true>, m, c->s, m, b.sx.p, b.sy.p, b.sz.p, c->P.kernel_support_radius,
           c->reach, c->dthr, b.sg, b.sx.p, b.sy.p, b.sz.p, off.p, (uint32_t *)nullptr, idx.p);
    LAUNCH(k_static_volume, m, c->s, m, c->K, c->P.boundary_rest_volume_weighting, b.sx.p,
           b.sy.p, b.sz.p, off.p, idx.p, b.svol.p);
    HIPCK(hipStreamSynchronize(c->s));
}

// B::find_boundary_samples (sph/mod.rs:707-712)
static void boundary_samples(sph_ctx *c) {
    const size_t n = c->n;
    const double range = c->P.kernel_support_radius;
    const double vmax = 4.0 / 3.0 * 3.14159265358979323846 * ((range * range) * range);
    for (auto &bp : c->bodies) {
        Body &b = *bp;
        if (b.kind == 0) {
            b.soff.reserve(n + 1);
            b.scnt.reserve(std::max<size_t>(n, 1));
            int bb[6];
            // the query cells may lie outside the sample grid: guard the union box
            HIPCK(hipMemcpyAsync(bb, c->bbox.p, sizeof bb, hipMemcpyDeviceToHost, c->s));
            HIPCK(hipStreamSynchronize(c->s));
            int lo[3], hi[3];
            for (int k = 0; k < 3; k++) {
                lo[k] = std::min(bb[k], b.sg.org[k]);
                hi[k] = std::max(bb[3 + k], b.sg.org[k] + b.sg.dim[k] - 1);
            }
            if (n) hash_guard(c, lo, hi, c->reach);
            LAUNCH(k_neighbors<false>, n, c->s, n, c->x.p, c->y.p, c->z.p, range, c->reach,
                   c->dthr, b.sg, b.sx.p, b.sy.p, b.sz.p, (const uint32_t *)nullptr, b.scnt.p,
                   (uint32_t *)nullptr);
            uint32_t total = scan_u32(c, b.scnt.p, n, b.soff.p, true);
            b.sidx.reserve(std::max<uint32_t>(total, 1));
            LAUNCH(k_neighbors<true>, n, c->s, n, c->x.p, c->y.p, c->z.p, range, c->reach,
                   c->dthr, b.sg, b.sx.p, b.sy.p, b.sz.p, b.soff.p, (uint32_t *)nullptr,
                   b.sidx.p);
        } else {
            b.flag.reserve(n);
            DBuf<double> *f[] = {&b.bx, &b.by, &b.bz, &b.bvx, &b.bvy, &b.bvz, &b.bvol};
            for (auto *q : f) q->reserve(n);
            PLAUNCH(c, SPH_PROF_BOUNDARY_SAMPLE, k_vm_sample, n, n, c->x.p, c->y.p, c->z.p, b.view(), range,
                   c->P.rest_density_grid_spacing, vmax, b.flag.p, b.bx.p, b.by.p, b.bz.p,
                   b.bvx.p, b.bvy.p, b.bvz.p, b.bvol.p);
        }
    }
    upload_bodies(c);
}

static bool any_dynamic(const sph_ctx *c) {
    for (auto &b : c->bodies)
        if (b->dynamic) return true;
    return false;
}

// reduce per-block body partials and add them to the rigid bodies
static void apply_body_partials(sph_ctx *c, size_t n) {
    if (!any_dynamic(c)) return;
    unsigned nbk = nblocks(n);
    if (nbk == 0) return;
    const int rows = kMaxBodies * 6;
    hipLaunchKernelGGL(k_reduce_rows, dim3(rows), dim3(kBlock), 0, c->s, c->partials.p,
                       (int)nbk, rows, c->red.p, 0);
    HIPCK(hipGetLastError());
    double out[kMaxBodies * 6];
    HIPCK(hipMemcpyAsync(out, c->red.p, sizeof out, hipMemcpyDeviceToHost, c->s));
    HIPCK(hipStreamSynchronize(c->s));
    for (size_t bi = 0; bi < c->bodies.size(); bi++) {
        Body &b = *c->bodies[bi];
        if (!b.dynamic) continue;
        b.rb.add_force_torque(out + bi * 6);
        for (int k = 0; k < 6; k++) b.applied[k] += out[bi * 6 + k];
    }
}

static void pressure_accel(sph_ctx *c, double *tx, double *ty, double *tz, int overwrite,
                           int mirror) {
    const size_t n = c->n;
    PLAUNCH(c, SPH_PROF_PRESSURE_ACCEL, k_pressure_accel, n, n, c->K, c->P.boundary_pressure_acceleration_weighting,
           c->x.p, c->y.p, c->z.p, c->m.p, c->V.p, c->p.p, ELL(c),
           c->d_bodies.p, (int)c->bodies.size(), overwrite, mirror, tx, ty, tz, c->partials.p);
    if (mirror) apply_body_partials(c, n);
}

static SolveConst solve_const(const sph_ctx *c) {
    SolveConst S;
    S.dt = c->dt;
    S.V0 = c->V0;
    S.eps = c->P.min_diagonal_element;
    S.omega = c->P.relaxation_factor;
    S.gamma = c->P.boundary_pressure_acceleration_weighting;
    return S;
}

// IISPH::resolve_pressure_globally (iisph.rs:378-560); after < 0: target
// density error, else AfterIteration(after)
static uint32_t resolve_pressure(sph_ctx *c, int vp, int after) {
    const size_t n = c->n;
    if (n == 0) return 0;
    SolveConst S = solve_const(c);
    const int nb = (int)c->bodies.size();
    PLAUNCH(c, SPH_PROF_SOURCE_DIAG, k_source_diag, n, n, c->K, S, vp, c->x.p, c->y.p, c->z.p, c->svx.p, c->svy.p,
           c->svz.p, c->m.p, c->V.p, ELL(c), c->d_bodies.p, nb, c->s_f.p,
           c->a_ff.p, c->p.p);
    uint32_t iter = 0;
    double err = INFINITY;
    c->err_hist.clear();
    const unsigned nbk = nblocks(n);
    for (;;) {
        bool conv = after >= 0 ? (iter > (uint32_t)after) : (iter > 2 && err < c->P.target_density_error);
        if (conv) break;
        pressure_accel(c, c->pax.p, c->pay.p, c->paz.p, 1, 0);
        PLAUNCH(c, SPH_PROF_JACOBI, k_jacobi, n, n, c->K, S, c->x.p, c->y.p, c->z.p, c->V.p, c->pax.p,
               c->pay.p, c->paz.p, c->s_f.p, c->a_ff.p, ELL(c),
               c->d_bodies.p, nb, c->p.p, c->partials.p);
        hipLaunchKernelGGL(k_reduce_rows, dim3(1), dim3(kBlock), 0, c->s, c->partials.p,
                           (int)nbk, 1, c->red.p, 0);
        HIPCK(hipGetLastError());
        HIPCK(hipMemcpyAsync(c->pinned, c->red.p, sizeof(double), hipMemcpyDeviceToHost, c->s));
        HIPCK(hipStreamSynchronize(c->s));
        if (c->prof.on) c->prof.drain();
        err = c->pinned[0] / (double)n * 100.0;
        c->err_hist.push_back(err);
        iter++;
        if (iter >= 100000u) throw Error(SPH_ERR_UNSUPPORTED, "pressure solve did not converge");
    }
    c->last_iter = iter;
    c->last_err = err;
    return iter;
}

static void solve_pressure(sph_ctx *c) {
    const size_t n = c->n;
    const double dt = c->dt;
    if (c->P.solver == SPH_SOLVER_SESPH) {  // sesph.rs:31-67
        LAUNCH(k_sesph_pressure, n, c->s, n, c->P.stiffness, c->V0, c->V.p, c->p.p);
        pressure_accel(c, c->ax.p, c->ay.p, c->az.p, 0, 1);
    } else if (c->P.solver == SPH_SOLVER_IISPH) {  // iisph.rs:55-96
        LAUNCH(k_pred_vel, n, c->s, n, dt, 0, c->vx.p, c->vy.p, c->vz.p, c->ax.p, c->ay.p,
               c->az.p, c->svx.p, c->svy.p, c->svz.p);
        resolve_pressure(c, 1, -1);
        pressure_accel(c, c->ax.p, c->ay.p, c->az.p, 0, 1);
    } else {  // iisph_optimized_source_term.rs:53-217
        LAUNCH(k_pred_vel, n, c->s, n, dt, 0, c->vx.p, c->vy.p, c->vz.p, c->ax.p, c->ay.p,
               c->az.p, c->svx.p, c->svy.p, c->svz.p);
        c->eqs1_iter = resolve_pressure(c, 0, 3);
        pressure_accel(c, c->ax.p, c->ay.p, c->az.p, 1, 1);
        LAUNCH(k_pred_vel, n, c->s, n, dt, 1, c->vx.p, c->vy.p, c->vz.p, c->ax.p, c->ay.p,
               c->az.p, c->svx.p, c->svy.p, c->svz.p);
        LAUNCH(k_pred_pos, n, c->s, n, dt, c->x.p, c->y.p, c->z.p, c->svx.p, c->svy.p, c->svz.p,
               c->sxp.p, c->syp.p, c->szp.p);
        resolve_pressure(c, 1, -1);
        pressure_accel(c, c->ax.p, c->ay.p, c->az.p, 1, 1);
        LAUNCH(k_ost_commit, n, c->s, n, dt, c->ax.p, c->ay.p, c->az.p, c->sxp.p, c->syp.p,
               c->szp.p, c->svx.p, c->svy.p, c->svz.p);
        // solver slots become the integrator slots (TakePredicted commits them)
        std::swap(c->ix.p, c->sxp.p); std::swap(c->ix.cap, c->sxp.cap);
        std::swap(c->iy.p, c->syp.p); std::swap(c->iy.cap, c->syp.cap);
        std::swap(c->iz.p, c->szp.p); std::swap(c->iz.cap, c->szp.cap);
        std::swap(c->ivx.p, c->svx.p); std::swap(c->ivx.cap, c->svx.cap);
        std::swap(c->ivy.p, c->svy.p); std::swap(c->ivy.cap, c->svy.cap);
        std::swap(c->ivz.p, c->svz.p); std::swap(c->ivz.cap, c->svz.cap);
    }
}

// reset_acceleration + gravity + viscosity (sph/mod.rs:640-657)
static void nonpressure(sph_ctx *c) {
    const size_t n = c->n;
    NonPressureConst NP;
    NP.gravity_mode = c->P.gravity_mode;
    for (int k = 0; k < 3; k++) NP.gcenter[k] = c->P.gravity_center[k];
    NP.nu_f = c->P.fluid_viscosity;
    NP.nu_b = c->P.boundary_viscosity;
    NP.eps2 = 0.01 * (c->P.rest_density_grid_spacing * c->P.rest_density_grid_spacing);
    // gravity onto dynamic bodies first (non_pressure_accelerations.rs:64-103)
    for (auto &bp : c->bodies) {
        Body &b = *bp;
        if (!b.dynamic) continue;
        double acc[3];
        if (c->P.gravity_mode == SPH_GRAVITY_UNIFORM) {
            acc[0] = 0.0;
            acc[1] = 0.0;
            acc[2] = -9.81;
        } else {
            double d[3] = {c->P.gravity_center[0] - b.rb.com[0], c->P.gravity_center[1] - b.rb.com[1],
                           c->P.gravity_center[2] - b.rb.com[2]};
            double dist = std::sqrt(((d[0] * d[0] + d[1] * d[1]) + d[2] * d[2]) + 1e-4);
            for (int k = 0; k < 3; k++) acc[k] = (9.81 * d[k]) / dist;
        }
        double f[3] = {b.rb.mass * acc[0], b.rb.mass * acc[1], b.rb.mass * acc[2]};
        b.rb.add_force(f, b.rb.com);
        for (int k = 0; k < 3; k++) b.applied[k] += f[k];
    }
    PLAUNCH(c, SPH_PROF_NONPRESSURE, k_nonpressure, n, n, c->K, NP, c->x.p, c->y.p, c->z.p, c->vx.p, c->vy.p,
           c->vz.p, c->m.p, c->V.p, ELL(c), c->d_bodies.p,
           (int)c->bodies.size(), c->ax.p, c->ay.p, c->az.p, c->partials.p);
    apply_body_partials(c, n);
}

static void measure_sums(sph_ctx *c, double out[3]) {
    const size_t n = c->n;
    if (n == 0) {
        out[0] = out[1] = out[2] = 0;
        return;
    }
    unsigned nbk = nblocks(n);
    LAUNCH(k_measure, n, c->s, n, c->V0, c->V.p, c->m.p, c->vx.p, c->vy.p, c->vz.p,
           c->partials.p);
    hipLaunchKernelGGL(k_reduce_rows, dim3(3), dim3(kBlock), 0, c->s, c->partials.p, (int)nbk, 3,
                       c->red.p, 0);
    HIPCK(hipGetLastError());
    HIPCK(hipMemcpyAsync(c->pinned, c->red.p, 3 * sizeof(double), hipMemcpyDeviceToHost, c->s));
    HIPCK(hipStreamSynchronize(c->s));
    for (int k = 0; k < 3; k++) out[k] = c->pinned[k];
}

// Fluid::disable/swap/drop_inactive on the reference index order (fluid.rs:159-210)
static void compact_below(sph_ctx *c) {
    const size_t n = c->n;
    std::vector<double> zh(n);
    std::vector<uint32_t> rh(n);
    c->z.download(zh.data(), n, c->s);
    c->ref.download(rh.data(), n, c->s);
    HIPCK(hipStreamSynchronize(c->s));
    std::vector<uint32_t> order(n);   // order[r] = storage slot of reference index r
    for (size_t k = 0; k < n; k++) order[rh[k]] = (uint32_t)k;
    size_t id = 0, nn = n;
    while (id < nn) {
        if (zh[order[id]] < c->P.disable_particles_below) {
            nn -= 1;
            std::swap(order[id], order[nn]);
        } else {
            id++;
        }
    }
    std::vector<uint32_t> newref(n, 0xffffffffu);
    for (size_t r = 0; r < nn; r++) newref[order[r]] = (uint32_t)r;
    std::vector<uint32_t> keep;
    keep.reserve(nn);
    std::vector<uint32_t> kref;
    kref.reserve(nn);
    for (size_t k = 0; k < n; k++)
        if (newref[k] != 0xffffffffu) {
            keep.push_back((uint32_t)k);
            kref.push_back(newref[k]);
        }
    c->perm.upload(keep.data(), nn, c->s);
    DBuf<double> *f64s[] = {&c->x, &c->y, &c->z, &c->vx, &c->vy, &c->vz, &c->m};
    for (auto *f : f64s) permute(c, *f, c->tmp, nn);
    permute(c, c->fid, c->tmpu, nn);
    c->ref.upload(kref.data(), nn, c->s);
    HIPCK(hipStreamSynchronize(c->s));
    c->n = nn;
}

// max |v| over the fluid and the number of particles below the disable threshold
static double speed_below(sph_ctx *c, double *vmax) {
    const size_t n = c->n;
    if (n == 0) {
        *vmax = 0.0;
        return 0.0;
    }
    unsigned nbk = nblocks(n);
    LAUNCH(k_speed_below, n, c->s, n, c->vx.p, c->vy.p, c->vz.p, c->z.p,
           c->P.disable_particles_below, c->partials.p);
    hipLaunchKernelGGL(k_fold_speed, dim3(1), dim3(kBlock), 0, c->s, c->partials.p, (int)nbk,
                       c->red.p);
    HIPCK(hipGetLastError());
    HIPCK(hipMemcpyAsync(c->pinned, c->red.p, 2 * sizeof(double), hipMemcpyDeviceToHost, c->s));
    HIPCK(hipStreamSynchronize(c->s));
    *vmax = c->pinned[0];
    return c->pinned[1];
}

// System::update (sph/mod.rs:674-731)
static void update(sph_ctx *c) {
    size_t n = c->n;
    double vmax = 0.0;
    double below = speed_below(c, &vmax);
    if (below > 0.0) {
        compact_below(c);
        n = c->n;
        speed_below(c, &vmax);  // max speed over the remaining particles
    }
    if (c->P.cfl) {  // set_cfl_time_step (sph/mod.rs:825-829)
        double cfl = c->P.cfl_number * c->P.rest_density_grid_spacing / vmax;
        c->dt = std::fmin(c->P.max_time_increment, cfl);
    }
    for (auto &b : c->bodies)
        for (int k = 0; k < 6; k++) b->applied[k] = 0;
    fluid_neighbors(c);
    boundary_samples(c);
    pairs_and_volume(c);
    nonpressure(c);
    solve_pressure(c);
    double sums[3];
    measure_sums(c, sums);
    c->avg_density = n ? sums[0] / (double)n : 0.0;
    if (c->prof.on) c->prof.drain();
}

static void integrate(sph_ctx *c) {
    const size_t n = c->n;
    if (c->P.integrator == SPH_INTEGRATOR_EULER_CROMER) {
        LAUNCH(k_integrate_ec, n, c->s, n, c->dt, c->x.p, c->y.p, c->z.p, c->vx.p, c->vy.p,
               c->vz.p, c->ax.p, c->ay.p, c->az.p);
    } else {  // take_predicted.rs:14-27
        std::swap(c->x.p, c->ix.p); std::swap(c->x.cap, c->ix.cap);
        std::swap(c->y.p, c->iy.p); std::swap(c->y.cap, c->iy.cap);
        std::swap(c->z.p, c->iz.p); std::swap(c->z.cap, c->iz.cap);
        std::swap(c->vx.p, c->ivx.p); std::swap(c->vx.cap, c->ivx.cap);
        std::swap(c->vy.p, c->ivy.p); std::swap(c->vy.cap, c->ivy.cap);
        std::swap(c->vz.p, c->ivz.p); std::swap(c->vz.cap, c->ivz.cap);
    }
}

static void fill_measurement(sph_ctx *c, sph_measurement *o) {
    double sums[3];
    measure_sums(c, sums);
    memset(o, 0, sizeof *o);
    const size_t n = c->n;
    o->time = c->time;
    o->time_step_size = c->dt;
    o->density = c->avg_density;
    o->density_error = n ? 100. * (sums[1] / (double)n - 1.) : 0.0;
    o->kinetic_energy = n ? sums[2] / (double)n : 0.0;
    o->solver_iterations = c->P.solver == SPH_SOLVER_SESPH ? 0 : c->last_iter;
    o->eqs1_iterations = c->P.solver == SPH_SOLVER_IISPHWOST ? c->eqs1_iter : 0;
    o->predicted_density_error = c->P.solver == SPH_SOLVER_SESPH ? 0.0 : c->last_err;
    o->steps = c->steps;
    o->num_particles = n;
}

static void validate_params(const sph_params &P) {
    if (!(P.kernel_support_radius > 0) || !(P.rest_density_grid_spacing > 0))
        throw Error(SPH_ERR_ARG, "kernel_support_radius and rest_density_grid_spacing must be > 0");
    if (P.solver < 0 || P.solver > 2) throw Error(SPH_ERR_ARG, "unknown solver");
    if (P.integrator < 0 || P.integrator > 1) throw Error(SPH_ERR_ARG, "unknown integrator");
    // PressureSolver::MANAGES_OWN_INTEGRATION pairing (setup/mod.rs:236-238)
    if ((P.solver == SPH_SOLVER_IISPHWOST) != (P.integrator == SPH_INTEGRATOR_TAKE_PREDICTED))
        throw Error(SPH_ERR_ARG, "IISPHwOST must be paired with TakePredicted and vice versa");
}

static void set_host_fluid(sph_ctx *c, size_t n, const double *xyz, const double *v,
                           const double *m, const uint32_t *fid) {
    if (n >= 0xffffffffull) throw Error(SPH_ERR_UNSUPPORTED, "more than 2^32-1 particles");
    reserve_fluid(c, n);
    std::vector<double> X(n), Y(n), Z(n);
    for (size_t i = 0; i < n; i++) {
        X[i] = xyz[3 * i];
        Y[i] = xyz[3 * i + 1];
        Z[i] = xyz[3 * i + 2];
    }
    c->x.upload(X.data(), n, c->s);
    c->y.upload(Y.data(), n, c->s);
    c->z.upload(Z.data(), n, c->s);
    for (size_t i = 0; i < n; i++) {
        X[i] = v ? v[3 * i] : 0.0;
        Y[i] = v ? v[3 * i + 1] : 0.0;
        Z[i] = v ? v[3 * i + 2] : 0.0;
    }
    c->vx.upload(X.data(), n, c->s);
    c->vy.upload(Y.data(), n, c->s);
    c->vz.upload(Z.data(), n, c->s);
    if (m) c->m.upload(m, n, c->s);
    else LAUNCH(k_fill_f64, n, c->s, c->m.p, n, 1.0);
    std::vector<uint32_t> ids(n), r(n);
    for (size_t i = 0; i < n; i++) {
        ids[i] = fid ? fid[i] : 0u;
        r[i] = (uint32_t)i;
    }
    c->fid.upload(ids.data(), n, c->s);
    c->ref.upload(r.data(), n, c->s);
    LAUNCH(k_fill_f64, n, c->s, c->ax.p, n, 0.0);
    LAUNCH(k_fill_f64, n, c->s, c->ay.p, n, 0.0);
    LAUNCH(k_fill_f64, n, c->s, c->az.p, n, 0.0);
    DBuf<double> *zero[] = {&c->ix, &c->iy, &c->iz, &c->ivx, &c->ivy, &c->ivz, &c->p, &c->V};
    for (auto *q : zero) LAUNCH(k_fill_f64, n, c->s, q->p, n, 0.0);
    HIPCK(hipStreamSynchronize(c->s));
    c->n = n;
}

// download a sorted-order f64 field into caller order (by reference index)
static void download_field(sph_ctx *c, const DBuf<double> &f, std::vector<double> &out) {
    const size_t n = c->n;
    std::vector<double> h(n);
    std::vector<uint32_t> r(n);
    f.download(h.data(), n, c->s);
    c->ref.download(r.data(), n, c->s);
    HIPCK(hipStreamSynchronize(c->s));
    out.assign(n, 0.0);
    for (size_t k = 0; k < n; k++) out[r[k]] = h[k];
}

static void init_boundaries(sph_ctx *c) {
    if (c->boundaries_initialized) return;
    for (auto &b : c->bodies)
        if (b->kind == 0) prepare_samples(c, *b);
    c->boundaries_initialized = true;
}

// prepare neighbors + boundary samples for a caller-provided state (per-trait
// entry points); the caller's index is the reference index
static void prepare_external(sph_ctx *c, size_t n, const double *x, const double *v,
                             const double *m) {
    init_boundaries(c);
    set_host_fluid(c, n, x, v, m, nullptr);
    fluid_neighbors(c);
    boundary_samples(c);
    pairs_and_volume(c);
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
    const size_t n = c->n;
    std::vector<uint32_t> r(n);
    c->ref.download(r.data(), n, c->s);
    HIPCK(hipStreamSynchronize(c->s));
    std::vector<double> h(n);
    for (size_t k = 0; k < n; k++) h[k] = src[stride * r[k] + comp];
    dst.upload(h.data(), n, c->s);
    HIPCK(hipStreamSynchronize(c->s));
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
        c->P = *params;
        const double h = params->kernel_support_radius, dxr = params->rest_density_grid_spacing;
        c->V0 = (dxr * dxr) * dxr;  // rest_volume = dx.powi(3) (sph/mod.rs:816)
        const double alpha = 16. / 3.14159265358979323846;
        c->K.h = h;
        c->K.wfac = alpha / ((h * h) * h);
        c->K.gfac = alpha / ((h * h) * (h * h));
        c->dthr = distance_threshold(h);
        c->reach = (int)std::ceil(h / h);
        c->dt = params->cfl ? 0.0 : params->time_increment;
        HIPCK(hipGetDevice(&c->device));
        HIPCK(hipStreamCreateWithFlags(&c->s, hipStreamNonBlocking));
        HIPCK(hipHostMalloc((void **)&c->pinned, 64 * sizeof(double), hipHostMallocDefault));
        reserve_fluid(c.get(), 1);
        *out = c.release();
    });
}

int sph_ctx_destroy(sph_ctx *c) {
    if (!c) return SPH_OK;
    return guarded([&] {
        if (c->s) (void)hipStreamSynchronize(c->s);
        if (c->pinned) (void)hipHostFree(c->pinned);
        hipStream_t s = c->s;
        delete c;
        if (s) (void)hipStreamDestroy(s);
    });
}

int sph_ctx_clone(const sph_ctx *src, sph_ctx **out) {
    return guarded([&] {
        if (!src || !out) throw Error(SPH_ERR_ARG, "null argument");
        sph_ctx *c = nullptr;
        int st = sph_ctx_create(&src->P, &c);
        if (st) throw Error(st, sph_last_error());
        std::unique_ptr<sph_ctx, int (*)(sph_ctx *)> guard(c, sph_ctx_destroy);
        sph_ctx *s = const_cast<sph_ctx *>(src);
        HIPCK(hipStreamSynchronize(s->s));
        const size_t n = s->n;
        reserve_fluid(c, n);
        DBuf<double> sph_ctx::*f64s[] = {
            &sph_ctx::x,   &sph_ctx::y,   &sph_ctx::z,   &sph_ctx::vx,  &sph_ctx::vy,
            &sph_ctx::vz,  &sph_ctx::ax,  &sph_ctx::ay,  &sph_ctx::az,  &sph_ctx::m,
            &sph_ctx::V,   &sph_ctx::p,   &sph_ctx::ix,  &sph_ctx::iy,  &sph_ctx::iz,
            &sph_ctx::ivx, &sph_ctx::ivy, &sph_ctx::ivz};
        for (auto f : f64s) (c->*f).copy_from(s->*f, n, c->s);
        c->fid.copy_from(s->fid, n, c->s);
        c->ref.copy_from(s->ref, n, c->s);
        c->n = n;
        for (auto &bp : s->bodies) {
            std::unique_ptr<Body> b(new Body());
            b->kind = bp->kind;
            b->dynamic = bp->dynamic;
            b->rb = bp->rb;
            b->ns = bp->ns;
            b->spos = bp->spos;
            memcpy(b->applied, bp->applied, sizeof b->applied);
            if (bp->kind == 1) {
                auto copy_lat = [&](LatticeHost &d, const LatticeHost &o) {
                    d.d = o.d;
                    size_t nodes = (3 * (size_t)o.d.n[0] + 1) * (3 * (size_t)o.d.n[1] + 1) *
                                   (3 * (size_t)o.d.n[2] + 1);
                    size_t cells = (size_t)o.d.n[0] * o.d.n[1] * o.d.n[2];
                    d.values.copy_from(o.values, nodes, c->s);
                    d.cell_ok.copy_from(o.cell_ok, cells, c->s);
                    d.d.values = d.values.p;
                    d.d.cell_ok = d.cell_ok.p;
                };
                copy_lat(b->sdf, bp->sdf);
                copy_lat(b->vm, bp->vm);
            }
            c->bodies.push_back(std::move(b));
        }
        c->boundaries_initialized = false;
        init_boundaries(c);
        c->dt = s->dt;
        c->time = s->time;
        c->steps = s->steps;
        c->last_iter = s->last_iter;
        c->eqs1_iter = s->eqs1_iter;
        c->last_err = s->last_err;
        c->avg_density = s->avg_density;
        c->err_hist = s->err_hist;
        HIPCK(hipStreamSynchronize(c->s));
        if (s->initialized) {
            // rebuild neighbor/boundary state so the clone is self-contained
            std::vector<double> keep_dt{c->dt};
            fluid_neighbors(c);
            boundary_samples(c);
            pairs_and_volume(c);
            c->dt = keep_dt[0];
        }
        c->initialized = s->initialized;
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
            X.upload(hx.data(), cnt, c->s);
            Y.upload(hy.data(), cnt, c->s);
            Z.upload(hz.data(), cnt, c->s);
            HIPCK(hipStreamSynchronize(c->s));
        };
        DBuf<double> qx, qy, qz, sx0, sy0, sz0, sx, sy, sz;
        soa(pos, n, qx, qy, qz);
        soa(samples, m, sx0, sy0, sz0);
        DBuf<int> sci, qci;
        sci.reserve(3 * std::max<size_t>(m, 1));
        qci.reserve(3 * std::max<size_t>(n, 1));
        DBuf<int> bbs, bbq;
        bbs.reserve(6);
        bbq.reserve(6);
        cell_box(c, m, cell_size, sx0.p, sy0.p, sz0.p, sci.p, bbs.p);
        cell_box(c, n, cell_size, qx.p, qy.p, qz.p, qci.p, bbq.p);
        int b1[6], b2[6];
        HIPCK(hipMemcpyAsync(b1, bbs.p, sizeof b1, hipMemcpyDeviceToHost, c->s));
        HIPCK(hipMemcpyAsync(b2, bbq.p, sizeof b2, hipMemcpyDeviceToHost, c->s));
        HIPCK(hipStreamSynchronize(c->s));
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
        int lo[3], hi[3];
        for (int k = 0; k < 3; k++) {
            lo[k] = std::min(b1[k], b2[k]);
            hi[k] = std::max(b1[3 + k], b2[3 + k]);
        }
        c->hash_box_valid = false;
        hash_guard(c, lo, hi, reach);
        c->hash_box_valid = false;
        GridDev g = make_grid(b1, b1 + 3);
        DBuf<uint32_t> cnt, start, perm;
        cell_sort(c, m, sci.p, g, cnt, start, perm, nullptr);
        g.start = start.p;
        sx.reserve(m);
        sy.reserve(m);
        sz.reserve(m);
        LAUNCH(k_gather<double>, m, c->s, m, perm.p, sx0.p, sx.p);
        LAUNCH(k_gather<double>, m, c->s, m, perm.p, sy0.p, sy.p);
        LAUNCH(k_gather<double>, m, c->s, m, perm.p, sz0.p, sz.p);
        DBuf<uint32_t> qcnt, off, idx;
        qcnt.reserve(n);
        off.reserve(n + 1);
        LAUNCH(k_neighbors<false>, n, c->s, n, qx.p, qy.p, qz.p, cell_size, reach, dthr, g, sx.p,
               sy.p, sz.p, (const uint32_t *)nullptr, qcnt.p, (uint32_t *)nullptr);
        uint32_t tot = scan_u32(c, qcnt.p, n, off.p, true);
        *total = tot;
        DBuf<uint64_t> off64;
        off64.reserve(n + 1);
        LAUNCH(k_widen_u32, n + 1, c->s, n + 1, off.p, off64.p);
        off64.download(offsets, n + 1, c->s);
        HIPCK(hipStreamSynchronize(c->s));
        if (tot > cap) throw Error(SPH_ERR_CAPACITY, "indices buffer too small");
        if (tot == 0) return;
        if (!indices) throw Error(SPH_ERR_ARG, "null indices");
        idx.reserve(tot);
        LAUNCH(k_neighbors<true>, n, c->s, n, qx.p, qy.p, qz.p, cell_size, reach, dthr, g, sx.p,
               sy.p, sz.p, off.p, (uint32_t *)nullptr, idx.p);
        DBuf<uint64_t> idx64;
        idx64.reserve(tot);
        LAUNCH(k_map_indices, tot, c->s, (size_t)tot, perm.p, idx.p, idx64.p);
        idx64.download(indices, tot, c->s);
        HIPCK(hipStreamSynchronize(c->s));
    });
}

int sph_boundary_add_samples(sph_ctx *c, const double *pos, size_t m, const sph_rigid_init *rigid) {
    int idx = -1;
    int st = guarded([&] {
        if (!c || (!pos && m)) throw Error(SPH_ERR_ARG, "null argument");
        if (c->bodies.size() >= (size_t)kMaxBodies) throw Error(SPH_ERR_UNSUPPORTED, "too many boundaries");
        if (rigid)
            throw Error(SPH_ERR_UNSUPPORTED,
                        "dynamic StaticSampleBoundary bodies are not built; use a volume map");
        std::unique_ptr<Body> b(new Body());
        b->kind = 0;
        b->ns = m;
        b->spos.assign(pos, pos + 3 * m);
        c->bodies.push_back(std::move(b));
        c->boundaries_initialized = false;
        idx = (int)c->bodies.size() - 1;
    });
    return st ? st : idx;
}

int sph_boundary_add_volume_map(sph_ctx *c, const sph_lattice *sdf, const sph_lattice *vm,
                                const sph_rigid_init *rigid) {
    int idx = -1;
    int st = guarded([&] {
        if (!c || !sdf || !vm || !sdf->values || !sdf->present || !vm->values || !vm->present)
            throw Error(SPH_ERR_ARG, "null argument");
        if (c->bodies.size() >= (size_t)kMaxBodies) throw Error(SPH_ERR_UNSUPPORTED, "too many boundaries");
        if (rigid && c->P.solver == SPH_SOLVER_IISPHWOST)  // SUPPORTS_DYNAMIC_BOUNDARIES (setup/mod.rs:242-244)
            throw Error(SPH_ERR_ARG, "IISPHwOST does not support dynamic boundaries");
        if (rigid && !(rigid->mass != 0.0)) throw Error(SPH_ERR_ARG, "rigid body mass cannot be zero");
        std::unique_ptr<Body> b(new Body());
        b->kind = 1;
        b->sdf.load(*sdf, c->s);
        b->vm.load(*vm, c->s);
        if (rigid) {
            b->dynamic = true;
            b->rb.init(*rigid);
        }
        c->bodies.push_back(std::move(b));
        idx = (int)c->bodies.size() - 1;
    });
    return st ? st : idx;
}

int sph_fluid_upload(sph_ctx *c, size_t n, const double *x, const double *v, const double *m,
                     const uint32_t *fluid_id) {
    return guarded([&] {
        if (!c || (!x && n) || (!m && n)) throw Error(SPH_ERR_ARG, "null argument");
        set_host_fluid(c, n, x, v, m, fluid_id);
        c->initialized = false;
    });
}

int sph_system_init(sph_ctx *c) {
    return guarded([&] {
        if (!c) throw Error(SPH_ERR_ARG, "null argument");
        init_boundaries(c);
        c->dt = c->P.cfl ? 0.0 : c->P.time_increment;
        c->time = 0;
        c->steps = 0;
        update(c);
        c->initialized = true;
    });
}

int sph_step(sph_ctx *c, int nsteps, sph_measurement *out) {
    return guarded([&] {
        if (!c) throw Error(SPH_ERR_ARG, "null argument");
        if (!c->initialized) throw Error(SPH_ERR_STATE, "sph_system_init has not run");
        for (int k = 0; k < nsteps; k++) {
            const double dt = c->dt;
            integrate(c);
            for (auto &b : c->bodies)
                if (b->dynamic) b->rb.step(dt);  // B::step_forward_in_time
            c->steps++;
            c->time += dt;
            update(c);
        }
        if (out) fill_measurement(c, out);
    });
}

int sph_fluid_len(const sph_ctx *c, size_t *n) {
    if (!c || !n) {
        set_error("null argument");
        return SPH_ERR_ARG;
    }
    *n = c->n;
    return SPH_OK;
}

int sph_fluid_download(sph_ctx *c, double *x, double *v, double *a, double *V, double *p,
                       double *m, uint32_t *fluid_id) {
    return guarded([&] {
        if (!c) throw Error(SPH_ERR_ARG, "null argument");
        std::vector<double> A, B, C;
        if (x) {
            download_field(c, c->x, A);
            download_field(c, c->y, B);
            download_field(c, c->z, C);
            split3(A, B, C, x);
        }
        if (v) {
            download_field(c, c->vx, A);
            download_field(c, c->vy, B);
            download_field(c, c->vz, C);
            split3(A, B, C, v);
        }
        if (a) {
            download_field(c, c->ax, A);
            download_field(c, c->ay, B);
            download_field(c, c->az, C);
            split3(A, B, C, a);
        }
        if (V) {
            download_field(c, c->V, A);
            memcpy(V, A.data(), A.size() * sizeof(double));
        }
        if (p) {
            download_field(c, c->p, A);
            memcpy(p, A.data(), A.size() * sizeof(double));
        }
        if (m) {
            download_field(c, c->m, A);
            memcpy(m, A.data(), A.size() * sizeof(double));
        }
        if (fluid_id) {
            const size_t n = c->n;
            std::vector<uint32_t> h(n), r(n);
            c->fid.download(h.data(), n, c->s);
            c->ref.download(r.data(), n, c->s);
            HIPCK(hipStreamSynchronize(c->s));
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
        *count = c->err_hist.size();
        if (out)
            for (size_t k = 0; k < std::min(cap, c->err_hist.size()); k++) out[k] = c->err_hist[k];
    });
}

int sph_body_state(sph_ctx *c, int bi, double out[22]) {
    return guarded([&] {
        if (!c || !out || bi < 0 || (size_t)bi >= c->bodies.size())
            throw Error(SPH_ERR_ARG, "bad boundary index");
        const Rigid &r = c->bodies[bi]->rb;
        memcpy(out, r.com, 3 * sizeof(double));
        memcpy(out + 3, r.q, 4 * sizeof(double));
        memcpy(out + 7, r.v, 3 * sizeof(double));
        memcpy(out + 10, r.L, 3 * sizeof(double));
        memcpy(out + 13, r.F, 3 * sizeof(double));
        memcpy(out + 16, r.T, 3 * sizeof(double));
        memcpy(out + 19, r.w, 3 * sizeof(double));
    });
}

int sph_body_forces(sph_ctx *c, int bi, double out[6]) {
    return guarded([&] {
        if (!c || !out || bi < 0 || (size_t)bi >= c->bodies.size())
            throw Error(SPH_ERR_ARG, "bad boundary index");
        memcpy(out, c->bodies[bi]->applied, 6 * sizeof(double));
    });
}

int sph_profile(sph_ctx *c, int enable) {
    return guarded([&] {
        if (!c) throw Error(SPH_ERR_ARG, "null argument");
        HIPCK(hipStreamSynchronize(c->s));
        c->prof.drain();
        c->prof.reset();
        c->prof.on = enable != 0;
    });
}

int sph_profile_read(sph_ctx *c, int kernel, uint64_t *launches, uint64_t *units, double *ms) {
    return guarded([&] {
        if (!c || kernel < 0 || kernel >= SPH_PROF_KERNELS) throw Error(SPH_ERR_ARG, "bad argument");
        HIPCK(hipStreamSynchronize(c->s));
        c->prof.drain();
        if (launches) *launches = c->prof.launches[kernel];
        if (units) *units = c->prof.units[kernel];
        if (ms) *ms = c->prof.ms[kernel];
    });
}

int sph_volume(sph_ctx *c, size_t n, const double *x, double *V_out) {
    return guarded([&] {
        if (!c || (!x && n) || (!V_out && n)) throw Error(SPH_ERR_ARG, "null argument");
        prepare_external(c, n, x, nullptr, nullptr);
        std::vector<double> A;
        download_field(c, c->V, A);
        if (n) memcpy(V_out, A.data(), n * sizeof(double));
        c->initialized = false;
    });
}

int sph_boundary_find_samples(sph_ctx *c, int bi, size_t n, const double *x,
                              double *samples_out, uint8_t *flags_out) {
    return guarded([&] {
        if (!c || (!x && n) || bi < 0 || (size_t)bi >= c->bodies.size())
            throw Error(SPH_ERR_ARG, "bad argument");
        Body &b = *c->bodies[bi];
        if (b.kind != 1) throw Error(SPH_ERR_ARG, "boundary is not a volume map");
        prepare_external(c, n, x, nullptr, nullptr);
        std::vector<double> f[7];
        const DBuf<double> *src[7] = {&b.bx, &b.by, &b.bz, &b.bvx, &b.bvy, &b.bvz, &b.bvol};
        for (int k = 0; k < 7; k++) download_field(c, *src[k], f[k]);
        std::vector<uint8_t> fl(n), flr(n);
        std::vector<uint32_t> r(n);
        b.flag.download(fl.data(), n, c->s);
        c->ref.download(r.data(), n, c->s);
        HIPCK(hipStreamSynchronize(c->s));
        for (size_t k = 0; k < n; k++) flr[r[k]] = fl[k];
        for (size_t i = 0; i < n; i++) {
            if (flags_out) flags_out[i] = flr[i];
            if (samples_out)
                for (int k = 0; k < 7; k++) samples_out[7 * i + k] = flr[i] ? f[k][i] : 0.0;
        }
        c->initialized = false;
    });
}

int sph_nonpressure_accel(sph_ctx *c, size_t n, const double *x, const double *v, const double *m,
                          const double *V, double *a_out, double *body_ft_out) {
    return guarded([&] {
        if (!c || (n && (!x || !v || !m || !V || !a_out))) throw Error(SPH_ERR_ARG, "null argument");
        prepare_external(c, n, x, v, m);
        upload_sorted(c, V, 1, 0, c->V);
        for (auto &b : c->bodies)
            for (int k = 0; k < 6; k++) b->applied[k] = 0;
        std::vector<Rigid> saved;
        for (auto &b : c->bodies) saved.push_back(b->rb);
        nonpressure(c);
        std::vector<double> A, B, C;
        download_field(c, c->ax, A);
        download_field(c, c->ay, B);
        download_field(c, c->az, C);
        split3(A, B, C, a_out);
        for (size_t bi = 0; bi < c->bodies.size(); bi++) {
            if (body_ft_out) memcpy(body_ft_out + 6 * bi, c->bodies[bi]->applied, 6 * sizeof(double));
            c->bodies[bi]->rb = saved[bi];  // stateless with respect to the bodies
        }
        c->initialized = false;
    });
}

int sph_pressure_solve(sph_ctx *c, size_t n, const double *x, const double *v, const double *m,
                       const double *V, double dt, double *a_inout, double *p_out,
                       double *body_ft_out, double *xpred_out, double *vpred_out,
                       sph_measurement *info_out) {
    return guarded([&] {
        if (!c || (n && (!x || !v || !m || !V || !a_inout))) throw Error(SPH_ERR_ARG, "null argument");
        prepare_external(c, n, x, v, m);
        upload_sorted(c, V, 1, 0, c->V);
        upload_sorted(c, a_inout, 3, 0, c->ax);
        upload_sorted(c, a_inout, 3, 1, c->ay);
        upload_sorted(c, a_inout, 3, 2, c->az);
        for (auto &b : c->bodies)
            for (int k = 0; k < 6; k++) b->applied[k] = 0;
        std::vector<Rigid> saved;
        for (auto &b : c->bodies) saved.push_back(b->rb);
        const double keep_dt = c->dt;
        c->dt = dt;
        solve_pressure(c);
        c->dt = keep_dt;
        std::vector<double> A, B, C;
        download_field(c, c->ax, A);
        download_field(c, c->ay, B);
        download_field(c, c->az, C);
        split3(A, B, C, a_inout);
        if (p_out) {
            download_field(c, c->p, A);
            memcpy(p_out, A.data(), n * sizeof(double));
        }
        if (c->P.solver == SPH_SOLVER_IISPHWOST) {
            if (xpred_out) {
                download_field(c, c->ix, A);
                download_field(c, c->iy, B);
                download_field(c, c->iz, C);
                split3(A, B, C, xpred_out);
            }
            if (vpred_out) {
                download_field(c, c->ivx, A);
                download_field(c, c->ivy, B);
                download_field(c, c->ivz, C);
                split3(A, B, C, vpred_out);
            }
        }
        for (size_t bi = 0; bi < c->bodies.size(); bi++) {
            if (body_ft_out) memcpy(body_ft_out + 6 * bi, c->bodies[bi]->applied, 6 * sizeof(double));
            c->bodies[bi]->rb = saved[bi];
        }
        if (info_out) {
            memset(info_out, 0, sizeof *info_out);
            info_out->solver_iterations = c->P.solver == SPH_SOLVER_SESPH ? 0 : c->last_iter;
            info_out->eqs1_iterations = c->P.solver == SPH_SOLVER_IISPHWOST ? c->eqs1_iter : 0;
            info_out->predicted_density_error = c->P.solver == SPH_SOLVER_SESPH ? 0.0 : c->last_err;
            info_out->time_step_size = dt;
            info_out->num_particles = n;
        }
        c->initialized = false;
    });
}

}  // extern "C"
