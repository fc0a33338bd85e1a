/*
 * refcpu.c — CPU oracle (f64) for the SPH particle hot path.
 *
 * TEST INFRASTRUCTURE ONLY: restates the reference solver
 * (SammyMcFly/fluid-simulation, simulation_lib) so that the HIP product can be
 * checked against it and timed beside it. Never linked into the product.
 *
 * Algorithmic choices deliberately follow the reference's CPU path, because
 * this library is also the timed CPU baseline:
 *   - hash-map grid of per-bucket index vectors, populated sequentially
 *     (spatial_hashing.rs:27-36);
 *   - per-particle neighbor vectors that are flattened afterwards
 *     (neighbor_search/mod.rs:118-155);
 *   - one loop per reference function, parallel over particles (OpenMP in
 *     place of rayon);
 *   - forces onto dynamic bodies collected per particle and applied
 *     sequentially in particle order (iteration.rs:171-191);
 *   - a fresh error buffer per Jacobi iteration (iisph.rs:441).
 * Reductions over particles are summed sequentially in particle order so the
 * oracle is deterministic and independent of the thread count.
 *
 * Compile with -ffp-contract=off (see oracle/Makefile).
 */
#include "refcpu.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define REF_PI 3.14159265358979323846
static const double ALPHA = 16. / REF_PI;
#define REF_MAX_SOLVER_ITERATIONS 10000u

/* ===================================================================== */
/* small vector helpers (nalgebra evaluation order)                        */
/* ===================================================================== */
static inline double dot3(const double *a, const double *b) {
    return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2];
}
static inline double norm_sq3(const double *a) { return (a[0] * a[0] + a[1] * a[1]) + a[2] * a[2]; }
static inline double norm3(const double *a) { return sqrt(norm_sq3(a)); }
static inline void sub3(const double *a, const double *b, double *o) {
    o[0] = a[0] - b[0];
    o[1] = a[1] - b[1];
    o[2] = a[2] - b[2];
}
static inline void cross3(const double *a, const double *b, double *o) {
    double x = a[1] * b[2] - a[2] * b[1];
    double y = a[2] * b[0] - a[0] * b[2];
    double z = a[0] * b[1] - a[1] * b[0];
    o[0] = x;
    o[1] = y;
    o[2] = z;
}
static inline void matvec3(const double *M, const double *v, double *o) {
    /* row-major M; nalgebra gemv sums columns in order */
    double x = (M[0] * v[0] + M[1] * v[1]) + M[2] * v[2];
    double y = (M[3] * v[0] + M[4] * v[1]) + M[5] * v[2];
    double z = (M[6] * v[0] + M[7] * v[1]) + M[8] * v[2];
    o[0] = x;
    o[1] = y;
    o[2] = z;
}
static void matmul3(const double *A, const double *B, double *C) {
    double T[9];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            T[i * 3 + j] = (A[i * 3 + 0] * B[0 * 3 + j] + A[i * 3 + 1] * B[1 * 3 + j]) +
                           A[i * 3 + 2] * B[2 * 3 + j];
    memcpy(C, T, sizeof T);
}
static void transpose3(const double *A, double *T) {
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) T[j * 3 + i] = A[i * 3 + j];
}

/* ===================================================================== */
/* kernel: kernel/mod.rs:22-46, cubic_spline.rs:8-17                      */
/* ===================================================================== */
static inline double powi3(double x) { return (x * x) * x; }
static inline double powi2(double x) { return x * x; }
static inline double powi4(double x) {
    double x2 = x * x;
    return x2 * x2;
}

/* w(q) = max(1-q,0)^3 - 4 max(1/2-q,0)^3   (cubic_spline.rs:12-13) */
double ref_w(double q) { return powi3(fmax(1. - q, 0.)) - 4. * powi3(fmax(1. / 2. - q, 0.)); }

/* dw/dq = -3 max(1-q,0)^2 + 12 max(1/2-q,0)^2   (cubic_spline.rs:15-16) */
double ref_dqw(double q) {
    return -3. * powi2(fmax(1. - q, 0.)) + 12. * powi2(fmax(1. / 2. - q, 0.));
}

/* W(r,h) = alpha / h^3 * w(|r|/h)   (kernel/mod.rs:29-31) */
double ref_kernel_function(const double r[3], double h) {
    return ALPHA / powi3(h) * ref_w(norm3(r) / h);
}

/* grad W = alpha / h^4 * r / |r| * w'(q), zero at r = 0   (kernel/mod.rs:37-45) */
void ref_kernel_gradient(const double r[3], double h, double out[3]) {
    double rn = norm3(r);
    if (rn == 0.0) {
        out[0] = out[1] = out[2] = 0.0;
        return;
    }
    double s = ALPHA / powi4(h);
    double d = ref_dqw(rn / h);
    for (int k = 0; k < 3; k++) out[k] = ((s * r[k]) / rn) * d;
}

/* ===================================================================== */
/* spatial hashing: spatial_hashing.rs                                      */
/* ===================================================================== */
static inline int32_t sat_i32(double v) {
    if (v != v) return 0;
    if (v >= 2147483647.0) return 2147483647;
    if (v <= -2147483648.0) return (int32_t)(-2147483647 - 1);
    return (int32_t)v;
}

/* get_cell: floor(x / cell_size) as i32   (spatial_hashing.rs:40-46) */
void ref_get_cell(const double p[3], double cs, int32_t out[3]) {
    for (int k = 0; k < 3; k++) out[k] = sat_i32(floor(p[k] / cs));
}

/* map_int_to_uint   (spatial_hashing.rs:49-55) */
uint64_t ref_map_int_to_uint(int32_t k) {
    if (k >= 0) return (uint64_t)k * 2u;
    return ((uint64_t)(-(int64_t)k)) * 2u + 1u;
}

/* hash(k,l,m) = u(k)*P1 ^ u(l)*P2 ^ u(m)*P3   (spatial_hashing.rs:61-71) */
uint64_t ref_hash(int32_t k, int32_t l, int32_t m) {
    const uint64_t P1 = 73856093u, P2 = 19349663u, P3 = 83492791u;
    return (ref_map_int_to_uint(k) * P1) ^ (ref_map_int_to_uint(l) * P2) ^
           (ref_map_int_to_uint(m) * P3);
}

/* growable index vector (Vec<usize>) */
typedef struct {
    size_t len, cap;
    uint64_t *v;
} uvec;
static inline void uvec_push(uvec *u, uint64_t x) {
    if (u->len == u->cap) {
        u->cap = u->cap ? 2 * u->cap : 4;
        u->v = (uint64_t *)realloc(u->v, u->cap * sizeof(uint64_t));
    }
    u->v[u->len++] = x;
}

/* FxHashMap<u64, Vec<usize>> stand-in: open addressing on the bucket key */
typedef struct {
    size_t cap;       /* power of two */
    uint64_t *keys;
    int64_t *slot;    /* -1 empty, else index into buckets */
    uvec *buckets;
    size_t nb;
} hmap;

static inline size_t mix64(uint64_t x) {
    x ^= x >> 33;
    x *= 0xff51afd7ed558ccdULL;
    x ^= x >> 33;
    return (size_t)x;
}

static void hmap_init(hmap *h, size_t expected) {
    size_t cap = 16;
    while (cap < 2 * expected + 16) cap <<= 1;
    h->cap = cap;
    h->keys = (uint64_t *)malloc(cap * sizeof(uint64_t));
    h->slot = (int64_t *)malloc(cap * sizeof(int64_t));
    for (size_t i = 0; i < cap; i++) h->slot[i] = -1;
    h->buckets = (uvec *)calloc(expected + 1, sizeof(uvec));
    h->nb = 0;
}
static void hmap_free(hmap *h) {
    for (size_t i = 0; i < h->nb; i++) free(h->buckets[i].v);
    free(h->buckets);
    free(h->keys);
    free(h->slot);
}
static inline const uvec *hmap_get(const hmap *h, uint64_t key) {
    size_t i = mix64(key) & (h->cap - 1);
    while (h->slot[i] >= 0) {
        if (h->keys[i] == key) return &h->buckets[h->slot[i]];
        i = (i + 1) & (h->cap - 1);
    }
    return NULL;
}
static inline uvec *hmap_entry(hmap *h, uint64_t key) {
    size_t i = mix64(key) & (h->cap - 1);
    while (h->slot[i] >= 0) {
        if (h->keys[i] == key) return &h->buckets[h->slot[i]];
        i = (i + 1) & (h->cap - 1);
    }
    h->keys[i] = key;
    h->slot[i] = (int64_t)h->nb;
    return &h->buckets[h->nb++];
}

/* SpatialHashing::find_samples   (spatial_hashing.rs:130-158 with populate :27-36,
 * get_particles_in_range :75-103, NeighborList::rebuild/flatten mod.rs:118-155) */
int ref_find_samples(double cs, double range, const double *pos, size_t n,
                     const double *samples, size_t m, ref_nlist *out) {
    hmap hm;
    hmap_init(&hm, m);
    for (size_t id = 0; id < m; id++) { /* sequential populate */
        int32_t c[3];
        ref_get_cell(samples + 3 * id, cs, c);
        uvec_push(hmap_entry(&hm, ref_hash(c[0], c[1], c[2])), id);
    }
    uvec *per = (uvec *)calloc(n ? n : 1, sizeof(uvec));
    int32_t mult = (int32_t)ceil(range / cs);
#pragma omp parallel for schedule(dynamic, 256)
    for (size_t i = 0; i < n; i++) {
        const double *p = pos + 3 * i;
        int32_t c[3];
        ref_get_cell(p, cs, c);
        for (int32_t dx = -mult; dx <= mult; dx++)
            for (int32_t dy = -mult; dy <= mult; dy++)
                for (int32_t dz = -mult; dz <= mult; dz++) {
                    const uvec *b = hmap_get(&hm, ref_hash(c[0] + dx, c[1] + dy, c[2] + dz));
                    if (!b) continue;
                    for (size_t k = 0; k < b->len; k++) {
                        uint64_t j = b->v[k];
                        double d[3];
                        sub3(p, samples + 3 * j, d); /* (position - sample).norm() */
                        if (norm3(d) < range) uvec_push(&per[i], j);
                    }
                }
    }
    size_t total = 0;
    for (size_t i = 0; i < n; i++) total += per[i].len;
    out->n = n;
    out->total = total;
    out->offsets = (uint64_t *)malloc((n + 1) * sizeof(uint64_t));
    out->indices = (uint64_t *)malloc((total ? total : 1) * sizeof(uint64_t));
    size_t o = 0;
    out->offsets[0] = 0;
    for (size_t i = 0; i < n; i++) {
        if (per[i].len) memcpy(out->indices + o, per[i].v, per[i].len * sizeof(uint64_t));
        o += per[i].len;
        out->offsets[i + 1] = o;
        free(per[i].v);
    }
    free(per);
    hmap_free(&hm);
    return 0;
}

void ref_nlist_free(ref_nlist *nl) {
    if (!nl) return;
    free(nl->offsets);
    free(nl->indices);
    nl->offsets = nl->indices = NULL;
    nl->n = nl->total = 0;
}
ref_nlist *ref_find_samples_alloc(double cs, double range, const double *pos, size_t n,
                                  const double *samples, size_t m) {
    ref_nlist *nl = (ref_nlist *)calloc(1, sizeof(ref_nlist));
    ref_find_samples(cs, range, pos, n, samples, m, nl);
    return nl;
}
size_t ref_nlist_total(const ref_nlist *nl) { return nl->total; }
void ref_nlist_copy(const ref_nlist *nl, uint64_t *offsets, uint64_t *indices) {
    memcpy(offsets, nl->offsets, (nl->n + 1) * sizeof(uint64_t));
    if (nl->total) memcpy(indices, nl->indices, nl->total * sizeof(uint64_t));
}
void ref_nlist_destroy(ref_nlist *nl) {
    ref_nlist_free(nl);
    free(nl);
}

/* ===================================================================== */
/* cubic serendipity: utilities/discretization.rs:263-557                  */
/* ===================================================================== */
/* reference_nodes   (discretization.rs:267-300) */
void ref_reference_nodes(double out[96]) {
    int k = 0;
    const double s[2] = {-1.0, 1.0};
    const double t[2] = {-1. / 3., 1. / 3.};
    for (int iz = 0; iz < 2; iz++)
        for (int iy = 0; iy < 2; iy++)
            for (int ix = 0; ix < 2; ix++) {
                out[k++] = s[ix];
                out[k++] = s[iy];
                out[k++] = s[iz];
            }
    for (int iz = 0; iz < 2; iz++)
        for (int iy = 0; iy < 2; iy++)
            for (int it = 0; it < 2; it++) {
                out[k++] = t[it];
                out[k++] = s[iy];
                out[k++] = s[iz];
            }
    for (int iz = 0; iz < 2; iz++)
        for (int ix = 0; ix < 2; ix++)
            for (int it = 0; it < 2; it++) {
                out[k++] = s[ix];
                out[k++] = t[it];
                out[k++] = s[iz];
            }
    for (int iy = 0; iy < 2; iy++)
        for (int ix = 0; ix < 2; ix++)
            for (int it = 0; it < 2; it++) {
                out[k++] = s[ix];
                out[k++] = s[iy];
                out[k++] = t[it];
            }
}

static int is_corner(double c) { return fabs(fabs(c) - 1.0) < 1e-9; }

/* shape_functions   (discretization.rs:306-348) */
void ref_shape_functions(double xi, double eta, double zeta, double out[32]) {
    double nodes[96];
    ref_reference_nodes(nodes);
    double r2 = (xi * xi + eta * eta) + zeta * zeta;
    for (int i = 0; i < 32; i++) {
        double px = nodes[3 * i], py = nodes[3 * i + 1], pz = nodes[3 * i + 2];
        int cx = is_corner(px), cy = is_corner(py), cz = is_corner(pz);
        double v;
        if (cx && cy && cz)
            v = (1.0 / 64.0) * (1.0 + xi * px) * (1.0 + eta * py) * (1.0 + zeta * pz) *
                (9.0 * r2 - 19.0);
        else if (!cx && cy && cz)
            v = (9.0 / 64.0) * (1.0 - xi * xi) * (1.0 + 9.0 * xi * px) * (1.0 + eta * py) *
                (1.0 + zeta * pz);
        else if (cx && !cy && cz)
            v = (9.0 / 64.0) * (1.0 - eta * eta) * (1.0 + 9.0 * eta * py) * (1.0 + xi * px) *
                (1.0 + zeta * pz);
        else if (cx && cy && !cz)
            v = (9.0 / 64.0) * (1.0 - zeta * zeta) * (1.0 + 9.0 * zeta * pz) * (1.0 + xi * px) *
                (1.0 + eta * py);
        else
            v = 0.0;
        out[i] = v;
    }
}

static int node_offset(double r) { return (int)round((r + 1.0) * 1.5); } /* to_offset :262-264 */

/* per-node lattice offsets within a cell, in reference_nodes order; equals
 * to_offset() applied to ref_reference_nodes() (checked by the oracle tests) */
static const int NODE_OFFS[32][3] = {
    {0, 0, 0}, {3, 0, 0}, {0, 3, 0}, {3, 3, 0}, {0, 0, 3}, {3, 0, 3}, {0, 3, 3}, {3, 3, 3},
    {1, 0, 0}, {2, 0, 0}, {1, 3, 0}, {2, 3, 0}, {1, 0, 3}, {2, 0, 3}, {1, 3, 3}, {2, 3, 3},
    {0, 1, 0}, {0, 2, 0}, {3, 1, 0}, {3, 2, 0}, {0, 1, 3}, {0, 2, 3}, {3, 1, 3}, {3, 2, 3},
    {0, 0, 1}, {0, 0, 2}, {3, 0, 1}, {3, 0, 2}, {0, 3, 1}, {0, 3, 2}, {3, 3, 1}, {3, 3, 2}};

int ref_node_offset(double r) { return node_offset(r); }

static inline size_t lat_index(const ref_lattice *l, int64_t ix, int64_t iy, int64_t iz) {
    int64_t nx = 3 * l->n[0] + 1, ny = 3 * l->n[1] + 1;
    return (size_t)((iz * ny + iy) * nx + ix);
}

/* function   (discretization.rs:437-483) */
int ref_serendipity_function(const ref_lattice *l, const double p[3], double *out) {
    int64_t c[3];
    for (int d = 0; d < 3; d++) {
        double fi = floor((p[d] - l->x_min[d]) / l->dx);
        if (p[d] < l->x_min[d] || p[d] > l->x_max[d]) return 1;
        int64_t idx = (int64_t)fi;
        if (idx > l->n[d] - 1) idx = l->n[d] - 1;
        c[d] = idx;
    }
    const int(*offs)[3] = NODE_OFFS;
    int64_t base[3] = {3 * c[0], 3 * c[1], 3 * c[2]};
    for (int i = 0; i < 32; i++)
        if (!l->present[lat_index(l, base[0] + offs[i][0], base[1] + offs[i][1],
                                  base[2] + offs[i][2])])
            return 2;
    double o[3];
    for (int d = 0; d < 3; d++) o[d] = l->x_min[d] + (double)c[d] * l->dx;
    double xi = 2.0 * (p[0] - o[0]) / l->dx - 1.0;
    double eta = 2.0 * (p[1] - o[1]) / l->dx - 1.0;
    double zeta = 2.0 * (p[2] - o[2]) / l->dx - 1.0;
    double shp[32];
    ref_shape_functions(xi, eta, zeta, shp);
    double s = -0.0;
    for (int i = 0; i < 32; i++)
        s += l->values[lat_index(l, base[0] + offs[i][0], base[1] + offs[i][1],
                                 base[2] + offs[i][2])] *
             shp[i];
    *out = s;
    return 0;
}

/* gradient by central differences, step dx/6   (discretization.rs:527-557) */
int ref_serendipity_gradient(const ref_lattice *l, const double p[3], double out[3]) {
    double h = l->dx / 6.0;
    for (int a = 0; a < 3; a++) {
        double pp[3] = {p[0], p[1], p[2]}, pm[3] = {p[0], p[1], p[2]};
        pp[a] += h;
        pm[a] -= h;
        double fp, fm;
        int ep = ref_serendipity_function(l, pp, &fp);
        int em = ref_serendipity_function(l, pm, &fm);
        if (ep) return ep;
        if (em) return em;
        out[a] = (fp - fm) / (2.0 * h);
    }
    return 0;
}

/* ===================================================================== */
/* rigid body: rigid_body_motion.rs                                          */
/* ===================================================================== */
/* UnitQuaternion::to_rotation_matrix (nalgebra) */
void ref_quat_to_matrix(const double q[4], double R[9]) {
    double w = q[0], i = q[1], j = q[2], k = q[3];
    double ww = w * w, ii = i * i, jj = j * j, kk = k * k;
    double ij = i * j * 2.0, wk = w * k * 2.0, wj = w * j * 2.0, ik = i * k * 2.0,
           jk = j * k * 2.0, wi = w * i * 2.0;
    R[0] = ww + ii - jj - kk;
    R[1] = ij - wk;
    R[2] = wj + ik;
    R[3] = wk + ij;
    R[4] = ww - ii + jj - kk;
    R[5] = jk - wi;
    R[6] = ik - wj;
    R[7] = wi + jk;
    R[8] = ww - ii - jj + kk;
}

/* UnitQuaternion * Vector3 (nalgebra): t = 2 (q_v x v); v' = w t + q_v x t + v */
void ref_quat_rotate(const double q[4], const double v[3], double out[3]) {
    double qv[3] = {q[1], q[2], q[3]};
    double t[3], c[3];
    cross3(qv, v, t);
    for (int k = 0; k < 3; k++) t[k] *= 2.0;
    cross3(qv, t, c);
    for (int k = 0; k < 3; k++) out[k] = t[k] * q[0] + c[k] + v[k];
}

static void update_derived(ref_rigid *rb) { /* rigid_body_motion.rs:71-76 */
    double R[9], RT[9], T[9];
    ref_quat_to_matrix(rb->q, R);
    transpose3(R, RT);
    matmul3(R, rb->inertia_inv_body, T);
    matmul3(T, RT, rb->inertia_inv_world);
    matvec3(rb->inertia_inv_world, rb->ang_mom, rb->ang_vel);
}

/* RigidBodyMotion::new   (rigid_body_motion.rs:38-62) */
void ref_rigid_init(ref_rigid *rb, double mass, const double Ib[9], const double Iib[9],
                    const double com[3], const double q[4], const double v[3],
                    const double w[3]) {
    memset(rb, 0, sizeof *rb);
    rb->mass = mass;
    memcpy(rb->inertia_inv_body, Iib, 9 * sizeof(double));
    memcpy(rb->com, com, 3 * sizeof(double));
    memcpy(rb->q, q, 4 * sizeof(double));
    memcpy(rb->lin_vel, v, 3 * sizeof(double));
    double R[9], RT[9], T[9], Iw[9];
    ref_quat_to_matrix(q, R);
    transpose3(R, RT);
    matmul3(R, Ib, T);
    matmul3(T, RT, Iw);
    matvec3(Iw, w, rb->ang_mom);
    update_derived(rb);
}

/* add_force: F += f; tau += (x - c) x f   (rigid_body_motion.rs:143-146) */
void ref_rigid_add_force(ref_rigid *rb, const double f[3], const double loc[3]) {
    double r[3], t[3];
    for (int k = 0; k < 3; k++) rb->force[k] += f[k];
    sub3(loc, rb->com, r);
    cross3(r, f, t);
    for (int k = 0; k < 3; k++) rb->torque[k] += t[k];
}

/* UnitQuaternion::from_scaled_axis = exp(v/2) with identity below eps^2 */
static void quat_from_scaled_axis(const double v[3], double q[4]) {
    double h[3] = {v[0] / 2.0, v[1] / 2.0, v[2] / 2.0};
    double nn = norm_sq3(h);
    const double eps = 2.220446049250313e-16;
    if (nn <= eps * eps) {
        q[0] = 1.0;
        q[1] = q[2] = q[3] = 0.0;
        return;
    }
    double n = sqrt(nn);
    double s = 1.0 * sin(n) / n; /* w_exp = exp(0) = 1 */
    q[0] = 1.0 * cos(n);
    q[1] = h[0] * s;
    q[2] = h[1] * s;
    q[3] = h[2] * s;
}

static void quat_mul(const double a[4], const double b[4], double o[4]) {
    double w = a[0] * b[0] - a[1] * b[1] - a[2] * b[2] - a[3] * b[3];
    double i = a[0] * b[1] + a[1] * b[0] + a[2] * b[3] - a[3] * b[2];
    double j = a[0] * b[2] - a[1] * b[3] + a[2] * b[0] + a[3] * b[1];
    double k = a[0] * b[3] + a[1] * b[2] - a[2] * b[1] + a[3] * b[0];
    o[0] = w;
    o[1] = i;
    o[2] = j;
    o[3] = k;
}

/* step_forward_in_time (Euler-Cromer)   (rigid_body_motion.rs:183-201) */
void ref_rigid_step(ref_rigid *rb, double dt) {
    for (int k = 0; k < 3; k++) {
        double acc = rb->force[k] / rb->mass;
        rb->lin_vel[k] += acc * dt;
        rb->com[k] += rb->lin_vel[k] * dt;
        rb->ang_mom[k] += rb->torque[k] * dt;
    }
    matvec3(rb->inertia_inv_world, rb->ang_mom, rb->ang_vel);
    double sa[3] = {rb->ang_vel[0] * dt, rb->ang_vel[1] * dt, rb->ang_vel[2] * dt};
    double dq[4], nq[4];
    quat_from_scaled_axis(sa, dq);
    quat_mul(dq, rb->q, nq);
    memcpy(rb->q, nq, sizeof nq);
    for (int k = 0; k < 3; k++) rb->force[k] = rb->torque[k] = 0.0;
    update_derived(rb);
}

void ref_rigid_world_to_local(const ref_rigid *rb, const double p[3], double out[3]) {
    double d[3];
    sub3(p, rb->com, d);
    double qc[4] = {rb->q[0], -rb->q[1], -rb->q[2], -rb->q[3]};
    ref_quat_rotate(qc, d, out);
}
void ref_rigid_local_to_world(const ref_rigid *rb, const double p[3], double out[3]) {
    double r[3];
    ref_quat_rotate(rb->q, p, r);
    for (int k = 0; k < 3; k++) out[k] = r[k] + rb->com[k];
}
void ref_rigid_rotate(const ref_rigid *rb, const double v[3], double out[3]) {
    ref_quat_rotate(rb->q, v, out);
}
/* velocity_at_point = v + omega x (p - c)   (rigid_body_motion.rs:132-137) */
void ref_rigid_velocity_at_point(const ref_rigid *rb, const double p[3], double out[3]) {
    double r[3], c[3];
    sub3(p, rb->com, r);
    cross3(rb->ang_vel, r, c);
    for (int k = 0; k < 3; k++) out[k] = rb->lin_vel[k] + c[k];
}

/* ===================================================================== */
/* system                                                                    */
/* ===================================================================== */
typedef struct ref_boundary {
    int kind; /* 0 static samples, 1 volume map */
    int dynamic;
    ref_rigid rb;
    /* static samples (static_sample_boundary.rs:285-310) */
    size_t ns;
    double *spos_body, *spos, *svel, *svol;
    ref_nlist nl;
    /* volume map (volume_map_boundary.rs:1750-1768) */
    ref_lattice sdf, vm;
    double *sdf_vals, *vm_vals;
    uint8_t *sdf_pres, *vm_pres;
    size_t n_normal, n_visc;
    double *epos, *evel, *evol; /* normal entries then viscosity entries */
    uint64_t *noff, *voff;      /* n + 1 */
    uint64_t *nid, *vid;        /* global entry ids */
    /* per-mode views: off[mode], ids[mode] */
    const uint64_t *off[2];
    const uint64_t *ids[2];
    const double *pos, *vel, *vol;
    double applied[6]; /* F, tau applied during the last update */
    /* body-frame bounding box of the boundary's mesh (LocalAABB,
     * volume_map_boundary.rs:1772-1776); has_aabb = 0 keeps the boundary
     * out of rigid-rigid contact */
    int has_aabb;
    double aabb_c[3], aabb_h[3];
} ref_boundary;

struct ref_system {
    ref_params P;
    double rest_volume;
    double dt;
    double time;
    uint64_t steps;
    /* fixed dt: time = time_offset + (steps - offset_steps) * dt (sph/mod.rs:107-116) */
    double time_offset;
    uint64_t offset_steps;
    size_t n, cap;
    double *x, *v, *a, *m, *V, *p;
    uint32_t *fid;
    uint64_t *origin;
    double *ix, *iv; /* integrator slots */
    double *sx, *sv; /* solver slots */
    double *s_f, *a_ff, *pacc;
    ref_nlist nl;
    int nb;
    ref_boundary *b;
    double avg_density;
    uint32_t last_iter, eqs1_iter;
    /* rigid-rigid contact: Gauss-Legendre rule of the contact-cell quadrature
     * (degree 0: off) and the last step's active cells and NNCG iterations */
    int contact_degree;
    double contact_x[32], contact_w[32];
    uint32_t contact_cells, contact_iters;
    double *contact_rec; /* REF_CONTACT_REC doubles per active cell of the last step */
    double last_err;
    double *err_hist;
    size_t err_hist_len, err_hist_cap;
};

static double *dalloc(size_t k) { return (double *)calloc(k ? k : 1, sizeof(double)); }

ref_system *ref_system_create(const ref_params *P, size_t n, const double *x, const double *v,
                              const double *m, const uint32_t *fid) {
    ref_system *s = (ref_system *)calloc(1, sizeof(ref_system));
    s->P = *P;
    s->rest_volume = powi3(P->rest_density_grid_spacing); /* sph/mod.rs:816 */
    s->dt = P->cfl ? 0.0 : P->time_increment;
    s->n = s->cap = n;
    s->x = dalloc(3 * n);
    s->v = dalloc(3 * n);
    s->a = dalloc(3 * n);
    s->m = dalloc(n);
    s->V = dalloc(n);
    s->p = dalloc(n);
    s->fid = (uint32_t *)calloc(n ? n : 1, sizeof(uint32_t));
    s->origin = (uint64_t *)calloc(n ? n : 1, sizeof(uint64_t));
    memcpy(s->x, x, 3 * n * sizeof(double));
    if (v) memcpy(s->v, v, 3 * n * sizeof(double));
    memcpy(s->m, m, n * sizeof(double));
    if (fid) memcpy(s->fid, fid, n * sizeof(uint32_t));
    for (size_t i = 0; i < n; i++) s->origin[i] = i;
    s->ix = dalloc(3 * n);
    s->iv = dalloc(3 * n);
    s->sx = dalloc(3 * n);
    s->sv = dalloc(3 * n);
    s->s_f = dalloc(n);
    s->a_ff = dalloc(n);
    s->pacc = dalloc(3 * n);
    s->err_hist_cap = 4096;
    s->err_hist = dalloc(s->err_hist_cap);
    return s;
}

void ref_system_destroy(ref_system *s) {
    if (!s) return;
    free(s->x); free(s->v); free(s->a); free(s->m); free(s->V); free(s->p);
    free(s->fid); free(s->origin);
    free(s->ix); free(s->iv); free(s->sx); free(s->sv);
    free(s->s_f); free(s->a_ff); free(s->pacc); free(s->err_hist);
    free(s->contact_rec);
    ref_nlist_free(&s->nl);
    for (int i = 0; i < s->nb; i++) {
        ref_boundary *b = &s->b[i];
        free(b->spos_body); free(b->spos); free(b->svel); free(b->svol);
        ref_nlist_free(&b->nl);
        free(b->sdf_vals); free(b->vm_vals); free(b->sdf_pres); free(b->vm_pres);
        free(b->epos); free(b->evel); free(b->evol);
        free(b->noff); free(b->voff); free(b->nid); free(b->vid);
    }
    free(s->b);
    free(s);
}

static ref_boundary *new_boundary(ref_system *s) {
    s->b = (ref_boundary *)realloc(s->b, (size_t)(s->nb + 1) * sizeof(ref_boundary));
    ref_boundary *b = &s->b[s->nb++];
    memset(b, 0, sizeof *b);
    return b;
}

static void rigid_from_init(ref_rigid *rb, const double *ri) {
    ref_rigid_init(rb, ri[0], ri + 1, ri + 10, ri + 19, ri + 22, ri + 26, ri + 29);
}

/* StaticSampleBoundary::update_positions_and_velocities (static_sample_boundary.rs:393-412) */
static void ss_update_positions(ref_boundary *b) {
    if (!b->dynamic) return;
    for (size_t k = 0; k < b->ns; k++) {
        ref_rigid_local_to_world(&b->rb, b->spos_body + 3 * k, b->spos + 3 * k);
        ref_rigid_velocity_at_point(&b->rb, b->spos + 3 * k, b->svel + 3 * k);
    }
}

int ref_system_add_static_samples(ref_system *s, const double *pos, size_t m,
                                  const double *rigid_init) {
    ref_boundary *b = new_boundary(s);
    b->kind = 0;
    b->ns = m;
    b->spos = dalloc(3 * m);
    b->svel = dalloc(3 * m);
    b->svol = dalloc(m);
    if (rigid_init) {
        b->dynamic = 1;
        rigid_from_init(&b->rb, rigid_init);
        b->spos_body = dalloc(3 * m);
        memcpy(b->spos_body, pos, 3 * m * sizeof(double));
    } else {
        memcpy(b->spos, pos, 3 * m * sizeof(double));
    }
    return s->nb - 1;
}

static void copy_lattice(ref_lattice *dst, const ref_lattice *src, double **vals,
                         uint8_t **pres) {
    *dst = *src;
    size_t cnt = (size_t)((3 * src->n[0] + 1) * (3 * src->n[1] + 1) * (3 * src->n[2] + 1));
    *vals = dalloc(cnt);
    *pres = (uint8_t *)malloc(cnt);
    memcpy(*vals, src->values, cnt * sizeof(double));
    memcpy(*pres, src->present, cnt);
    dst->values = *vals;
    dst->present = *pres;
}

int ref_system_add_volume_map(ref_system *s, const ref_lattice *sdf, const ref_lattice *vm,
                              const double *rigid_init) {
    ref_boundary *b = new_boundary(s);
    b->kind = 1;
    copy_lattice(&b->sdf, sdf, &b->sdf_vals, &b->sdf_pres);
    copy_lattice(&b->vm, vm, &b->vm_vals, &b->vm_pres);
    if (rigid_init) {
        b->dynamic = 1;
        rigid_from_init(&b->rb, rigid_init);
    }
    return s->nb - 1;
}

/* StaticSample BoundaryType::initialize: pseudo volumes   (static_sample_boundary.rs:354-390) */
static void ss_initialize(ref_system *s, ref_boundary *b) {
    ss_update_positions(b);
    ref_nlist bb;
    double h = s->P.kernel_support_radius;
    ref_find_samples(h, h, b->spos, b->ns, b->spos, b->ns, &bb);
    for (size_t k = 0; k < b->ns; k++) {
        double inv = 0.;
        for (uint64_t e = bb.offsets[k]; e < bb.offsets[k + 1]; e++) {
            double r[3];
            sub3(b->spos + 3 * bb.indices[e], b->spos + 3 * k, r);
            inv += ref_kernel_function(r, h);
        }
        b->svol[k] = s->P.boundary_rest_volume_weighting / inv;
    }
    ref_nlist_free(&bb);
}

/* ---------------------------------------------------------------------- */
/* boundary sampling                                                        */
/* ---------------------------------------------------------------------- */
static void ss_find_samples(ref_system *s, ref_boundary *b) { /* :157-168 */
    ref_nlist_free(&b->nl);
    double h = s->P.kernel_support_radius;
    ref_find_samples(h, h, s->x, s->n, b->spos, b->ns, &b->nl);
    b->off[0] = b->off[1] = b->nl.offsets;
    b->ids[0] = b->ids[1] = b->nl.indices;
    b->pos = b->spos;
    b->vel = b->svel;
    b->vol = b->svol;
}

static int vm_sd(const ref_boundary *b, const double *p, double *out) {
    if (!b->dynamic) return ref_serendipity_function(&b->sdf, p, out);
    double pl[3];
    ref_rigid_world_to_local(&b->rb, p, pl);
    return ref_serendipity_function(&b->sdf, pl, out);
}
static int vm_sdg(const ref_boundary *b, const double *p, double out[3]) {
    if (!b->dynamic) return ref_serendipity_gradient(&b->sdf, p, out);
    double pl[3], gl[3];
    ref_rigid_world_to_local(&b->rb, p, pl);
    int e = ref_serendipity_gradient(&b->sdf, pl, gl);
    if (e) return e;
    ref_rigid_rotate(&b->rb, gl, out);
    return 0;
}
static int vm_vol(const ref_boundary *b, const double *p, double *out) {
    if (!b->dynamic) return ref_serendipity_function(&b->vm, p, out);
    double pl[3];
    ref_rigid_world_to_local(&b->rb, p, pl);
    return ref_serendipity_function(&b->vm, pl, out);
}
static void vm_vel(const ref_boundary *b, const double *p, double out[3]) {
    if (!b->dynamic) {
        out[0] = out[1] = out[2] = 0.0;
        return;
    }
    ref_rigid_velocity_at_point(&b->rb, p, out);
}

/* per-particle record: 1 normal sample + 4 viscosity samples */
typedef struct {
    int ok;
    double pos[3], vel[3], vol;
    double vpos[4][3], vvel[4][3], vvol;
} vm_rec;

/* VolumeMapBoundary::find_boundary_samples   (volume_map_boundary.rs:1259-1424) */
static void vm_find_samples(ref_system *s, ref_boundary *b) {
    size_t n = s->n;
    double range = s->P.kernel_support_radius, dxr = s->P.rest_density_grid_spacing;
    vm_rec *r = (vm_rec *)calloc(n ? n : 1, sizeof(vm_rec));
    double vmax = 4.0 / 3.0 * REF_PI * powi3(range); /* ball_volume :34-36 */
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < n; i++) {
        const double *pos = s->x + 3 * i;
        vm_rec *q = &r[i];
        q->ok = 0;
        double sd;
        if (vm_sd(b, pos, &sd) || !(sd < range)) continue;
        double g[3];
        if (vm_sdg(b, pos, g)) continue;
        double gn = norm3(g);
        if (!(gn > 0.5 && gn < 2.0)) continue;
        double vol;
        if (vm_vol(b, pos, &vol) || !(vol > 0.)) continue;
        vol = fmin(vol, vmax);
        double d = fmax(sd + 0.25 * dxr, dxr);
        double pob[3];
        for (int k = 0; k < 3; k++) pob[k] = pos[k] - (g[k] / gn) * d;
        q->ok = 1;
        memcpy(q->pos, pob, sizeof pob);
        vm_vel(b, pob, q->vel);
        q->vol = vol;
        double vp[3];
        sub3(pob, pos, vp);
        double vpn = norm3(vp);
        double tmp[3] = {1., 0., 0.};
        double u[3] = {vp[0] / vpn, vp[1] / vpn, vp[2] / vpn};
        if (fabs(dot3(u, tmp)) > 0.9) {
            tmp[0] = 0.;
            tmp[1] = 1.;
        }
        double t1[3], t2[3];
        cross3(vp, tmp, t1);
        cross3(vp, t1, t2);
        double n1 = norm3(t1), n2 = norm3(t2);
        for (int k = 0; k < 3; k++) {
            t1[k] = t1[k] / n1;
            t2[k] = t2[k] / n2;
        }
        double dd = 0.5 * range;
        for (int k = 0; k < 3; k++) {
            q->vpos[0][k] = pob[k] + dd * t1[k];
            q->vpos[1][k] = pob[k] - dd * t1[k];
            q->vpos[2][k] = pob[k] + dd * t2[k];
            q->vpos[3][k] = pob[k] - dd * t2[k];
        }
        for (int j = 0; j < 4; j++) vm_vel(b, q->vpos[j], q->vvel[j]);
        q->vvol = 0.25 * vol;
    }
    /* sequential write-back into the two lists (global offset split) */
    size_t nn = 0;
    for (size_t i = 0; i < n; i++) nn += r[i].ok;
    free(b->epos); free(b->evel); free(b->evol);
    free(b->noff); free(b->voff); free(b->nid); free(b->vid);
    b->n_normal = nn;
    b->n_visc = 4 * nn;
    size_t tot = nn * 5;
    b->epos = dalloc(3 * tot);
    b->evel = dalloc(3 * tot);
    b->evol = dalloc(tot);
    b->noff = (uint64_t *)malloc((n + 1) * sizeof(uint64_t));
    b->voff = (uint64_t *)malloc((n + 1) * sizeof(uint64_t));
    b->nid = (uint64_t *)malloc((nn ? nn : 1) * sizeof(uint64_t));
    b->vid = (uint64_t *)malloc((nn ? 4 * nn : 1) * sizeof(uint64_t));
    size_t k = 0;
    b->noff[0] = b->voff[0] = 0;
    for (size_t i = 0; i < n; i++) {
        if (r[i].ok) {
            memcpy(b->epos + 3 * k, r[i].pos, 3 * sizeof(double));
            memcpy(b->evel + 3 * k, r[i].vel, 3 * sizeof(double));
            b->evol[k] = r[i].vol;
            b->nid[k] = k;
            for (int j = 0; j < 4; j++) {
                size_t e = nn + 4 * k + (size_t)j;
                memcpy(b->epos + 3 * e, r[i].vpos[j], 3 * sizeof(double));
                memcpy(b->evel + 3 * e, r[i].vvel[j], 3 * sizeof(double));
                b->evol[e] = r[i].vvol;
                b->vid[4 * k + (size_t)j] = e;
            }
            k++;
        }
        b->noff[i + 1] = k;
        b->voff[i + 1] = 4 * k;
    }
    b->off[0] = b->noff;
    b->off[1] = b->voff;
    b->ids[0] = b->nid;
    b->ids[1] = b->vid;
    b->pos = b->epos;
    b->vel = b->evel;
    b->vol = b->evol;
    free(r);
}

static void find_boundary_samples(ref_system *s) {
    for (int i = 0; i < s->nb; i++) {
        if (s->b[i].kind == 0)
            ss_find_samples(s, &s->b[i]);
        else
            vm_find_samples(s, &s->b[i]);
    }
}

/* ---------------------------------------------------------------------- */
/* per-particle loops                                                       */
/* ---------------------------------------------------------------------- */
/* get_volume   (quantities.rs:14-61) */
static void get_volume(ref_system *s) {
    double h = s->P.kernel_support_radius, V0 = s->rest_volume;
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < s->n; i++) {
        double accu = 0.;
        for (uint64_t e = s->nl.offsets[i]; e < s->nl.offsets[i + 1]; e++) {
            double r[3];
            sub3(s->x + 3 * i, s->x + 3 * s->nl.indices[e], r);
            accu += V0 * ref_kernel_function(r, h);
        }
        for (int bi = 0; bi < s->nb; bi++) {
            const ref_boundary *b = &s->b[bi];
            for (uint64_t e = b->off[0][i]; e < b->off[0][i + 1]; e++) {
                uint64_t id = b->ids[0][e];
                double r[3];
                sub3(s->x + 3 * i, b->pos + 3 * id, r);
                accu += b->vol[id] * ref_kernel_function(r, h);
            }
        }
        s->V[i] = V0 / accu;
    }
}

/* gravity   (non_pressure_accelerations.rs:64-103) */
static void add_gravity(ref_system *s) {
    const double g = 9.81;
    if (s->P.gravity_mode == REF_GRAVITY_UNIFORM) {
        for (size_t i = 0; i < s->n; i++) {
            s->a[3 * i + 0] += 0.0;
            s->a[3 * i + 1] += 0.0;
            s->a[3 * i + 2] += -g;
        }
        for (int bi = 0; bi < s->nb; bi++) {
            ref_boundary *b = &s->b[bi];
            if (!b->dynamic) continue;
            double acc[3] = {0.0, 0.0, -g}, f[3];
            for (int k = 0; k < 3; k++) f[k] = b->rb.mass * acc[k];
            ref_rigid_add_force(&b->rb, f, b->rb.com);
            for (int k = 0; k < 3; k++) b->applied[k] += f[k];
        }
    } else {
        const double *c = s->P.gravity_center;
        for (size_t i = 0; i < s->n; i++) {
            double dir[3];
            sub3(c, s->x + 3 * i, dir);
            double dist = sqrt(norm_sq3(dir) + 1e-4);
            for (int k = 0; k < 3; k++) s->a[3 * i + k] += (g * dir[k]) / dist;
        }
        for (int bi = 0; bi < s->nb; bi++) {
            ref_boundary *b = &s->b[bi];
            if (!b->dynamic) continue;
            double dir[3];
            sub3(c, b->rb.com, dir);
            double dist = sqrt(norm_sq3(dir) + 1e-4);
            double acc[3], f[3];
            for (int k = 0; k < 3; k++) acc[k] = (g * dir[k]) / dist;
            for (int k = 0; k < 3; k++) f[k] = b->rb.mass * acc[k];
            ref_rigid_add_force(&b->rb, f, b->rb.com);
            for (int k = 0; k < 3; k++) b->applied[k] += f[k];
        }
    }
}

/* collected force onto boundary */
typedef struct {
    int b;
    double f[3], loc[3];
} cforce;
typedef struct {
    size_t len, cap;
    cforce *v;
} cfvec;
static void cf_push(cfvec *c, int b, const double *f, const double *loc) {
    if (c->len == c->cap) {
        c->cap = c->cap ? 2 * c->cap : 8;
        c->v = (cforce *)realloc(c->v, c->cap * sizeof(cforce));
    }
    cforce *x = &c->v[c->len++];
    x->b = b;
    memcpy(x->f, f, 3 * sizeof(double));
    memcpy(x->loc, loc, 3 * sizeof(double));
}
/* apply collected forces sequentially in particle order (mod.rs:190-194) */
static void apply_forces(ref_system *s, cfvec *per, size_t n) {
    for (size_t i = 0; i < n; i++) {
        for (size_t k = 0; k < per[i].len; k++) {
            cforce *c = &per[i].v[k];
            ref_boundary *b = &s->b[c->b];
            if (!b->dynamic) continue;
            ref_rigid_add_force(&b->rb, c->f, c->loc);
            double r[3], t[3];
            sub3(c->loc, b->rb.com, r);
            cross3(r, c->f, t);
            for (int q = 0; q < 3; q++) {
                b->applied[q] += c->f[q];
                b->applied[3 + q] += t[q];
            }
        }
        free(per[i].v);
    }
    free(per);
}

/* add_viscosity_acceleration   (non_pressure_accelerations.rs:106-185) */
static void add_viscosity(ref_system *s) {
    double h = s->P.kernel_support_radius;
    double eps2 = 0.01 * powi2(s->P.rest_density_grid_spacing);
    double nu_f = s->P.fluid_viscosity, nu_b = s->P.boundary_viscosity;
    cfvec *per = (cfvec *)calloc(s->n ? s->n : 1, sizeof(cfvec));
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < s->n; i++) {
        const double *xi = s->x + 3 * i, *vi = s->v + 3 * i;
        double accu[3] = {0., 0., 0.};
        for (uint64_t e = s->nl.offsets[i]; e < s->nl.offsets[i + 1]; e++) {
            uint64_t j = s->nl.indices[e];
            double r[3], dv[3], gw[3];
            sub3(xi, s->x + 3 * j, r);
            sub3(vi, s->v + 3 * j, dv);
            ref_kernel_gradient(r, h, gw);
            double c = nu_f * 2. * (3. + 2.) * s->V[j] * dot3(dv, r) / (norm_sq3(r) + eps2);
            for (int k = 0; k < 3; k++) accu[k] += c * gw[k];
        }
        for (int bi = 0; bi < s->nb; bi++) {
            const ref_boundary *b = &s->b[bi];
            for (uint64_t e = b->off[1][i]; e < b->off[1][i + 1]; e++) {
                uint64_t id = b->ids[1][e];
                double r[3], dv[3], gw[3], acc[3];
                sub3(xi, b->pos + 3 * id, r);
                sub3(vi, b->vel + 3 * id, dv);
                ref_kernel_gradient(r, h, gw);
                double c = nu_b * 2. * (3. + 2.) * b->vol[id] * dot3(dv, r) / (norm_sq3(r) + eps2);
                for (int k = 0; k < 3; k++) acc[k] = c * gw[k];
                if (b->dynamic) {
                    double f[3];
                    for (int k = 0; k < 3; k++) f[k] = -(s->m[i] * acc[k]);
                    cf_push(&per[i], bi, f, b->pos + 3 * id);
                }
                for (int k = 0; k < 3; k++) accu[k] += acc[k];
            }
        }
        for (int k = 0; k < 3; k++) s->a[3 * i + k] += accu[k];
    }
    apply_forces(s, per, s->n);
}

/* add_pressure_acceleration   (pressure_solver/mod.rs:121-195) */
static void add_pressure_acceleration(ref_system *s, double *target, int overwrite,
                                      int mirror) {
    double h = s->P.kernel_support_radius, w = s->P.boundary_pressure_acceleration_weighting;
    cfvec *per = mirror ? (cfvec *)calloc(s->n ? s->n : 1, sizeof(cfvec)) : NULL;
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < s->n; i++) {
        const double *xi = s->x + 3 * i;
        double accu[3] = {0., 0., 0.};
        double Vi = s->V[i], mi = s->m[i], pi = s->p[i];
        for (uint64_t e = s->nl.offsets[i]; e < s->nl.offsets[i + 1]; e++) {
            uint64_t j = s->nl.indices[e];
            double r[3], gw[3];
            sub3(xi, s->x + 3 * j, r);
            ref_kernel_gradient(r, h, gw);
            double c = Vi / mi * s->V[j] * (pi + s->p[j]);
            for (int k = 0; k < 3; k++) accu[k] -= c * gw[k];
        }
        for (int bi = 0; bi < s->nb; bi++) {
            const ref_boundary *b = &s->b[bi];
            for (uint64_t e = b->off[0][i]; e < b->off[0][i + 1]; e++) {
                uint64_t id = b->ids[0][e];
                double r[3], gw[3], f[3];
                sub3(xi, b->pos + 3 * id, r);
                ref_kernel_gradient(r, h, gw);
                double c = 2. * w * Vi * b->vol[id] * pi;
                for (int k = 0; k < 3; k++) f[k] = c * gw[k];
                if (mirror && b->dynamic) cf_push(&per[i], bi, f, b->pos + 3 * id);
                for (int k = 0; k < 3; k++) accu[k] -= f[k] / mi;
            }
        }
        for (int k = 0; k < 3; k++) {
            if (overwrite)
                target[3 * i + k] = accu[k];
            else
                target[3 * i + k] += accu[k];
        }
    }
    if (mirror) apply_forces(s, per, s->n);
}

/* set_pred_vel_by_applying_acc   (pressure_solver/mod.rs:90-108) */
static void set_pred_vel(ref_system *s, int to_pred) {
    for (size_t i = 0; i < 3 * s->n; i++) {
        double base = to_pred ? s->sv[i] : s->v[i];
        s->sv[i] = base + s->dt * s->a[i];
    }
}

/* SESPH::solve_and_add_acceleration   (sesph.rs:31-67) */
static void sesph_solve(ref_system *s) {
    for (size_t i = 0; i < s->n; i++)
        s->p[i] = s->P.stiffness * fmax(s->rest_volume / s->V[i] - 1., 0.);
    add_pressure_acceleration(s, s->a, 0, 1);
}

/* SESPHwSplitting::solve_and_add_acceleration   (sesph_with_splitting.rs:31-81,
 * calc_predicted_density :97-153): v* = v + dt a, the density predicted from
 * the fluid's masses and v* (boundary samples mirror the particle's mass and
 * enter with v*_i alone), the state equation on m / rho*, then the pressure
 * acceleration as in SESPH */
static void sesph_split_solve(ref_system *s) {
    const double h = s->P.kernel_support_radius, dt = s->dt;
    set_pred_vel(s, 0);
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < s->n; i++) {
        const double *xi = s->x + 3 * i, *vi = s->sv + 3 * i;
        double accu = 0.;
        for (uint64_t e = s->nl.offsets[i]; e < s->nl.offsets[i + 1]; e++) {
            uint64_t j = s->nl.indices[e];
            double r[3], gw[3], dv[3];
            sub3(xi, s->x + 3 * j, r);
            sub3(vi, s->sv + 3 * j, dv);
            ref_kernel_gradient(r, h, gw);
            accu += s->m[j] * ref_kernel_function(r, h) + dt * dot3(dv, gw);
        }
        for (int bi = 0; bi < s->nb; bi++) {
            const ref_boundary *b = &s->b[bi];
            for (uint64_t e = b->off[0][i]; e < b->off[0][i + 1]; e++) {
                uint64_t id = b->ids[0][e];
                double r[3], gw[3];
                sub3(xi, b->pos + 3 * id, r);
                ref_kernel_gradient(r, h, gw);
                accu += b->vol[id] * s->m[i] / s->rest_volume * ref_kernel_function(r, h) +
                        dt * dot3(vi, gw);
            }
        }
        const double vol = s->m[i] / accu;
        s->p[i] = s->P.stiffness * fmax(s->rest_volume / vol - 1., 0.);
    }
    add_pressure_acceleration(s, s->a, 0, 1);
}

/* IISPH source terms   (iisph.rs:117-172 vde, 175-251 vp) */
static void set_source_term(ref_system *s, int vp) {
    double h = s->P.kernel_support_radius, dt = s->dt;
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < s->n; i++) {
        const double *xi = s->x + 3 * i, *vi = s->sv + 3 * i;
        double accu = vp ? 1. - s->rest_volume / s->V[i] : 0.;
        for (uint64_t e = s->nl.offsets[i]; e < s->nl.offsets[i + 1]; e++) {
            uint64_t j = s->nl.indices[e];
            double r[3], gw[3], dv[3];
            sub3(xi, s->x + 3 * j, r);
            sub3(vi, s->sv + 3 * j, dv);
            ref_kernel_gradient(r, h, gw);
            accu -= dt * s->V[j] * dot3(dv, gw);
        }
        for (int bi = 0; bi < s->nb; bi++) {
            const ref_boundary *b = &s->b[bi];
            for (uint64_t e = b->off[0][i]; e < b->off[0][i + 1]; e++) {
                uint64_t id = b->ids[0][e];
                double r[3], gw[3], dv[3];
                sub3(xi, b->pos + 3 * id, r);
                sub3(vi, b->vel + 3 * id, dv);
                ref_kernel_gradient(r, h, gw);
                accu -= dt * b->vol[id] * dot3(dv, gw);
            }
        }
        s->s_f[i] = accu;
    }
}

/* IISPH::set_diagonal_element   (iisph.rs:254-342) */
static void set_diagonal(ref_system *s) {
    double h = s->P.kernel_support_radius, dt = s->dt,
           w = s->P.boundary_pressure_acceleration_weighting;
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < s->n; i++) {
        const double *xi = s->x + 3 * i;
        double sf[3] = {0., 0., 0.}, sf2 = 0., sb[3] = {0., 0., 0.};
        double Vi = s->V[i];
        for (uint64_t e = s->nl.offsets[i]; e < s->nl.offsets[i + 1]; e++) {
            uint64_t j = s->nl.indices[e];
            double r[3], gw[3];
            sub3(xi, s->x + 3 * j, r);
            ref_kernel_gradient(r, h, gw);
            for (int k = 0; k < 3; k++) sf[k] += s->V[j] * gw[k];
            sf2 -= powi2(dt) * Vi * powi2(s->V[j]) / s->m[j] * norm_sq3(gw);
        }
        for (int bi = 0; bi < s->nb; bi++) {
            const ref_boundary *b = &s->b[bi];
            for (uint64_t e = b->off[0][i]; e < b->off[0][i + 1]; e++) {
                uint64_t id = b->ids[0][e];
                double r[3], gw[3];
                sub3(xi, b->pos + 3 * id, r);
                ref_kernel_gradient(r, h, gw);
                for (int k = 0; k < 3; k++) sb[k] += b->vol[id] * gw[k];
            }
        }
        double cf[3], ss[3];
        double ci = -Vi / s->m[i];
        for (int k = 0; k < 3; k++) {
            cf[k] = ci * (sf[k] + 2. * w * sb[k]);
            ss[k] = sf[k] + sb[k];
        }
        s->a_ff[i] = powi2(dt) * dot3(cf, ss) + sf2;
    }
}

/* IISPH::initialize   (iisph.rs:344-376), clamp_pressure = true */
static void iisph_initialize(ref_system *s) {
    double eps = s->P.min_diagonal_element, om = s->P.relaxation_factor;
    for (size_t i = 0; i < s->n; i++) {
        double aii = s->a_ff[i];
        if (aii > eps || aii < -eps)
            s->p[i] = fmax(om * s->s_f[i] / aii, 0.);
        else
            s->p[i] = 0.;
    }
}

/* IISPH::resolve_pressure_globally   (iisph.rs:378-560), clamp_pressure = true.
 * after_iteration < 0: TargetDensityError, else AfterIteration(n). */
static uint32_t resolve_pressure(ref_system *s, int after_iteration) {
    if (s->n == 0) return 0;
    double h = s->P.kernel_support_radius, dt = s->dt;
    double eps = s->P.min_diagonal_element, om = s->P.relaxation_factor;
    double tde = s->P.target_density_error;
    set_diagonal(s);
    iisph_initialize(s);
    uint32_t iter = 0;
    double err = INFINITY;
    s->err_hist_len = 0;
    for (;;) {
        int conv = after_iteration >= 0 ? (iter > (uint32_t)after_iteration)
                                        : (iter > 2 && err < tde);
        /* the reference has no cap (iisph.rs:387-390); this guard only stops a
         * diverged (NaN) or stalled solve from hanging the checker */
        if (conv || iter >= REF_MAX_SOLVER_ITERATIONS) break;
        add_pressure_acceleration(s, s->pacc, 1, 0);
        double *errs = dalloc(s->n); /* fresh buffer per iteration, as the reference */
#pragma omp parallel for schedule(static)
        for (size_t i = 0; i < s->n; i++) {
            const double *xi = s->x + 3 * i, *api = s->pacc + 3 * i;
            double adp = 0.;
            for (uint64_t e = s->nl.offsets[i]; e < s->nl.offsets[i + 1]; e++) {
                uint64_t j = s->nl.indices[e];
                double r[3], gw[3], d[3];
                sub3(xi, s->x + 3 * j, r);
                sub3(api, s->pacc + 3 * j, d);
                ref_kernel_gradient(r, h, gw);
                adp += powi2(dt) * s->V[j] * dot3(d, gw);
            }
            for (int bi = 0; bi < s->nb; bi++) {
                const ref_boundary *b = &s->b[bi];
                for (uint64_t e = b->off[0][i]; e < b->off[0][i + 1]; e++) {
                    uint64_t id = b->ids[0][e];
                    double r[3], gw[3];
                    sub3(xi, b->pos + 3 * id, r);
                    ref_kernel_gradient(r, h, gw);
                    adp += powi2(dt) * b->vol[id] * dot3(api, gw);
                }
            }
            if (s->a_ff[i] < -eps) {
                double pn = s->p[i] + om * (s->s_f[i] - adp) / s->a_ff[i];
                s->p[i] = fmax(pn, 0.);
            }
            errs[i] = (s->s_f[i] < 0.) ? fabs(adp - s->s_f[i]) : 0.;
        }
        double total = 0.;
        for (size_t i = 0; i < s->n; i++) total += errs[i];
        free(errs);
        err = total / (double)s->n * 100.0;
        if (s->err_hist_len < s->err_hist_cap) s->err_hist[s->err_hist_len++] = err;
        iter++;
    }
    s->last_iter = iter;
    s->last_err = err;
    return iter;
}

/* IISPH::solve_and_add_acceleration   (iisph.rs:55-96) */
static void iisph_solve(ref_system *s) {
    set_pred_vel(s, 0);
    set_source_term(s, 1);
    resolve_pressure(s, -1);
    add_pressure_acceleration(s, s->a, 0, 1);
}

/* IISPHwOST::solve_and_add_acceleration   (iisph_optimized_source_term.rs:53-217) */
static void iisphwost_solve(ref_system *s) {
    double dt = s->dt;
    set_pred_vel(s, 0);
    set_source_term(s, 0);
    s->eqs1_iter = resolve_pressure(s, 3);
    add_pressure_acceleration(s, s->a, 1, 1);
    set_pred_vel(s, 1);
    for (size_t i = 0; i < 3 * s->n; i++) s->sx[i] = s->x[i] + dt * s->sv[i];
    set_source_term(s, 1);
    resolve_pressure(s, -1);
    add_pressure_acceleration(s, s->a, 1, 1);
    /* new position and DFSPH-style velocity; the velocity Jacobian computed by the
     * reference at :155-183 never reaches any output and is not restated */
    for (size_t i = 0; i < 3 * s->n; i++) {
        double np = s->sx[i] + powi2(dt) * s->a[i];
        s->pacc[i] = s->sv[i] + dt * s->a[i];
        s->sx[i] = np;
    }
    memcpy(s->sv, s->pacc, 3 * s->n * sizeof(double));
    double *t;
    t = s->ix; s->ix = s->sx; s->sx = t;
    t = s->iv; s->iv = s->sv; s->sv = t;
}

/* calc_acceleration   (sph/mod.rs:658-671) */
static void calc_acceleration(ref_system *s) {
    memset(s->a, 0, 3 * s->n * sizeof(double));
    add_gravity(s);
    add_viscosity(s);
    if (s->P.solver == REF_SESPH)
        sesph_solve(s);
    else if (s->P.solver == REF_SESPH_SPLITTING)
        sesph_split_solve(s);
    else if (s->P.solver == REF_IISPH)
        iisph_solve(s);
    else
        iisphwost_solve(s);
}

/* Fluid::disable / swap   (fluid.rs:159-188) */
static void swap_particles(ref_system *s, size_t a, size_t b) {
    if (a == b) return;
#define SW3(arr)                                   \
    for (int k = 0; k < 3; k++) {                  \
        double t = arr[3 * a + k];                 \
        arr[3 * a + k] = arr[3 * b + k];           \
        arr[3 * b + k] = t;                        \
    }
#define SW1(arr, T)        \
    {                      \
        T t = arr[a];      \
        arr[a] = arr[b];   \
        arr[b] = t;        \
    }
    SW1(s->fid, uint32_t);
    SW3(s->x);
    SW3(s->v);
    SW3(s->a);
    SW1(s->m, double);
    SW1(s->V, double);
    SW1(s->p, double);
    SW3(s->ix);
    SW3(s->iv);
    SW3(s->sx);
    SW3(s->sv);
    SW1(s->origin, uint64_t);
#undef SW3
#undef SW1
}

static double calc_max_speed(const ref_system *s) { /* sph/mod.rs:424-441 */
    double mx = 0.0;
    for (size_t i = 0; i < s->n; i++) mx = fmax(mx, norm3(s->v + 3 * i));
    return mx;
}

static double calc_average_mass_density(const ref_system *s) { /* sph/mod.rs:444-486 */
    double sum = 0.;
    for (size_t i = 0; i < s->n; i++)
        sum += (s->V[i] < s->rest_volume) ? s->m[i] / s->V[i] : s->m[i] / s->rest_volume;
    return s->n ? sum / (double)s->n : 0.0;
}

/* System::update   (sph/mod.rs:674-731) */
static void update(ref_system *s) {
    size_t id = 0;
    while (id < s->n) {
        if (s->x[3 * id + 2] < s->P.disable_particles_below) {
            s->n -= 1;
            swap_particles(s, id, s->n);
        } else {
            id++;
        }
    }
    if (s->P.cfl) { /* set_cfl_time_step   (sph/mod.rs:825-829) */
        double vmax = calc_max_speed(s);
        double c = s->P.cfl_number * s->P.rest_density_grid_spacing / vmax;
        s->dt = fmin(s->P.max_time_increment, c);
    }
    double h = s->P.kernel_support_radius;
    ref_nlist_free(&s->nl);
    ref_find_samples(h, h, s->x, s->n, s->x, s->n, &s->nl);
    for (int i = 0; i < s->nb; i++) memset(s->b[i].applied, 0, sizeof s->b[i].applied);
    find_boundary_samples(s);
    get_volume(s);
    calc_acceleration(s);
    s->avg_density = calc_average_mass_density(s);
}

int ref_system_init(ref_system *s) {
    for (int i = 0; i < s->nb; i++)
        if (s->b[i].kind == 0) ss_initialize(s, &s->b[i]);
    update(s);
    return 0;
}

/* ===================================================================== */
/* rigid-rigid contact: VolumeMapBoundary::step_forward_in_time            */
/* (volume_map_boundary.rs:56-1102, 1440-1453)                             */
/* ===================================================================== */
#define CONTACT_CELL_SIZE_FACTOR 0.5
#define CONTACT_FRICTION 0.5
#define CONTACT_RELAXATION 0.5
#define CONTACT_MIN_DIAGONAL 1e-35
#define CONTACT_FB_TOL_PRESSURE 1e-6
#define CONTACT_FB_TOL_FRICTION 1e-3
#define CONTACT_MAX_ITERATIONS 100000u /* the reference has none (it warns at 100) */

/* BoundaryType::world_aabb (:2122-2147): |R| applied to the half extents */
static void contact_world_aabb(const ref_boundary *b, double mn[3], double mx[3]) {
    if (!b->dynamic) {
        for (int k = 0; k < 3; k++) {
            mn[k] = b->aabb_c[k] - b->aabb_h[k];
            mx[k] = b->aabb_c[k] + b->aabb_h[k];
        }
        return;
    }
    double R[9], hw[3], c[3];
    ref_quat_to_matrix(b->rb.q, R);
    for (int k = 0; k < 9; k++) R[k] = fabs(R[k]);
    matvec3(R, b->aabb_h, hw);
    ref_rigid_local_to_world(&b->rb, b->aabb_c, c); /* pose.transform_point */
    for (int k = 0; k < 3; k++) {
        mn[k] = c[k] - hw[k];
        mx[k] = c[k] + hw[k];
    }
}

static double ref_cubic_extension(double sd, double h);

/* cubic_extension_fn_derivative (:1734-1741) */
static double ref_cubic_extension_derivative(double sd, double h) {
    if (sd >= h || sd <= 0.0) return 0.0;
    double r[3] = {sd, 0., 0.}, z[3] = {0., 0., 0.}, g[3];
    ref_kernel_gradient(r, h, g);
    return g[0] / ref_kernel_function(z, h);
}

typedef struct contact_qp { /* QuadPointData (:832-842) */
    double pos[3], weight, ga[3], gb[3], sa, sb, spa, spb;
} contact_qp;

typedef struct contact_cell { /* ContactCell (:686-700) */
    int a, b;
    contact_qp *qp;
    size_t nqp;
    double volume, center[3], fa[3], taua[3], fb[3], taub[3], T[9], acc_p, acc_f;
} contact_cell;

/* build_contact_cell (:702-768): the degree^3 Gauss-Legendre points of the
 * sub-cell (:854-880) where both SDFs and gradients evaluate and both cubic
 * extensions of 2 phi + h are positive */
static int contact_build_cell(const ref_system *s, int ia, int ib, const double from[3],
                              const double to[3], contact_cell *cell) {
    const ref_boundary *A = &s->b[ia], *Bd = &s->b[ib];
    const double h = s->P.kernel_support_radius, scaling = 2.0;
    const int n = s->contact_degree;
    double half[3], center[3];
    for (int k = 0; k < 3; k++) half[k] = (to[k] - from[k]) * 0.5;
    for (int k = 0; k < 3; k++) center[k] = from[k] + half[k];
    double vc = 0.0;
    cell->qp = (contact_qp *)malloc((size_t)n * n * n * sizeof(contact_qp));
    cell->nqp = 0;
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++)
            for (int k = 0; k < n; k++) {
                double pos[3] = {center[0] + s->contact_x[i] * half[0],
                                 center[1] + s->contact_x[j] * half[1],
                                 center[2] + s->contact_x[k] * half[2]};
                double w = s->contact_w[i] * s->contact_w[j] * s->contact_w[k] * half[0] *
                           half[1] * half[2];
                double pa, pb, ga[3], gb[3];
                if (vm_sd(A, pos, &pa) || vm_sd(Bd, pos, &pb)) continue;
                double sa = ref_cubic_extension(scaling * pa + h, h);
                double sb = ref_cubic_extension(scaling * pb + h, h);
                if (sa <= 0.0 || sb <= 0.0) continue;
                if (vm_sdg(A, pos, ga) || vm_sdg(Bd, pos, gb)) continue;
                vc += w * sa * sb;
                contact_qp *q = &cell->qp[cell->nqp++];
                memcpy(q->pos, pos, sizeof pos);
                q->weight = w;
                memcpy(q->ga, ga, sizeof ga);
                memcpy(q->gb, gb, sizeof gb);
                q->sa = sa;
                q->sb = sb;
                q->spa = scaling * ref_cubic_extension_derivative(scaling * pa + h, h);
                q->spb = scaling * ref_cubic_extension_derivative(scaling * pb + h, h);
            }
    if (vc <= 0.0 || cell->nqp == 0) {
        free(cell->qp);
        cell->qp = NULL;
        return 0;
    }
    cell->a = ia;
    cell->b = ib;
    cell->volume = vc;
    return 1;
}

/* generate_contact_cells (:98-161) with broad_phase (:70-92) and
 * subdivide_overlap (:1813-1844): cells in the order of the pairs (i < j, one
 * of them dynamic, overlapping world AABBs) and of the sub-cells (z, y, x) */
typedef struct contact_task { /* one candidate sub-cell of a body pair */
    int a, b;
    double from[3], to[3];
} contact_task;

static contact_cell *contact_generate(const ref_system *s, size_t *ncell) {
    const double h = s->P.kernel_support_radius;
    size_t tcap = 64, nt = 0;
    contact_task *tasks = (contact_task *)malloc(tcap * sizeof(contact_task));
    for (int i = 0; i < s->nb; i++)
        for (int j = i + 1; j < s->nb; j++) {
            const ref_boundary *A = &s->b[i], *Bd = &s->b[j];
            if (!A->has_aabb || !Bd->has_aabb) continue;
            if (!A->dynamic && !Bd->dynamic) continue;
            double amn[3], amx[3], bmn[3], bmx[3];
            contact_world_aabb(A, amn, amx);
            contact_world_aabb(Bd, bmn, bmx);
            if (!(amn[0] <= bmx[0] && amx[0] >= bmn[0] && amn[1] <= bmx[1] &&
                  amx[1] >= bmn[1] && amn[2] <= bmx[2] && amx[2] >= bmn[2]))
                continue;
            double omn[3], omx[3];
            for (int k = 0; k < 3; k++) {
                omn[k] = fmax(amn[k], bmn[k]);
                omx[k] = fmin(amx[k], bmx[k]);
            }
            if (!(omn[0] <= omx[0] && omn[1] <= omx[1] && omn[2] <= omx[2])) continue;
            const double target = h * CONTACT_CELL_SIZE_FACTOR;
            double ext[3], cs[3];
            size_t nn[3];
            for (int k = 0; k < 3; k++) {
                ext[k] = omx[k] - omn[k];
                double c = ceil(ext[k] / target);
                size_t ck = c > 0.0 ? (size_t)c : 0; /* f64 as usize saturates at 0 */
                nn[k] = ck > 1 ? ck : 1;
                cs[k] = ext[k] / (double)nn[k];
            }
            for (size_t iz = 0; iz < nn[2]; iz++)
                for (size_t iy = 0; iy < nn[1]; iy++)
                    for (size_t ix = 0; ix < nn[0]; ix++) {
                        if (nt == tcap) {
                            tcap *= 2;
                            tasks = (contact_task *)realloc(tasks, tcap * sizeof(contact_task));
                        }
                        contact_task *t = &tasks[nt++];
                        t->a = i;
                        t->b = j;
                        t->from[0] = omn[0] + (double)ix * cs[0];
                        t->from[1] = omn[1] + (double)iy * cs[1];
                        t->from[2] = omn[2] + (double)iz * cs[2];
                        for (int k = 0; k < 3; k++) t->to[k] = t->from[k] + cs[k];
                    }
        }
    /* the candidate test (:124-146) and build_contact_cell per task; kept
     * cells in task order */
    contact_cell *tmp = (contact_cell *)calloc(nt ? nt : 1, sizeof(contact_cell));
    unsigned char *keep = (unsigned char *)calloc(nt ? nt : 1, 1);
#pragma omp parallel for schedule(dynamic, 16)
    for (size_t t = 0; t < nt; t++) {
        const contact_task *tk = &tasks[t];
        const ref_boundary *A = &s->b[tk->a], *Bd = &s->b[tk->b];
        double d[3], center[3], diff[3];
        for (int k = 0; k < 3; k++) d[k] = (tk->to[k] - tk->from[k]) * 0.5;
        for (int k = 0; k < 3; k++) center[k] = tk->from[k] + d[k];
        sub3(tk->to, tk->from, diff);
        const double bound = 2.0 * (norm3(diff) * 0.5);
        double pa, pb;
        if (vm_sd(A, center, &pa) || vm_sd(Bd, center, &pb)) continue;
        if (!(pa - bound < 0. && pb - bound < 0.)) continue;
        keep[t] = (unsigned char)contact_build_cell(s, tk->a, tk->b, tk->from, tk->to, &tmp[t]);
    }
    size_t nc = 0;
    for (size_t t = 0; t < nt; t++)
        if (keep[t]) tmp[nc++] = tmp[t];
    free(keep);
    free(tasks);
    *ncell = nc;
    return tmp;
}

/* compute_diagonal (:884-946) */
static void contact_diagonal(const ref_system *s, contact_cell *c, double dt) {
    const ref_boundary *A = &s->b[c->a], *Bd = &s->b[c->b];
    double x[3] = {0., 0., 0.}, fa[3] = {0., 0., 0.}, ta[3] = {0., 0., 0.},
           fb[3] = {0., 0., 0.}, tb[3] = {0., 0., 0.};
    for (size_t q = 0; q < c->nqp; q++) {
        const contact_qp *p = &c->qp[q];
        double wg = p->weight * p->sa * p->sb;
        for (int k = 0; k < 3; k++) x[k] += wg * p->pos[k];
        double wa = p->weight * p->spa * p->sb, r[3], cr[3];
        for (int k = 0; k < 3; k++) fa[k] += wa * p->ga[k];
        if (A->dynamic) {
            sub3(p->pos, A->rb.com, r);
            cross3(r, p->ga, cr);
            for (int k = 0; k < 3; k++) ta[k] += wa * cr[k];
        }
        double wb = p->weight * p->spb * p->sa;
        for (int k = 0; k < 3; k++) fb[k] += wb * p->gb[k];
        if (Bd->dynamic) {
            sub3(p->pos, Bd->rb.com, r);
            cross3(r, p->gb, cr);
            for (int k = 0; k < 3; k++) tb[k] += wb * cr[k];
        }
    }
    for (int k = 0; k < 3; k++) {
        c->center[k] = x[k] / c->volume;
        c->fa[k] = fa[k];
        c->taua[k] = ta[k];
        c->fb[k] = fb[k];
        c->taub[k] = tb[k];
    }
    const double fn = fmax(norm3(fa), 1e-15);
    double nrm[3] = {fa[0] / fn, fa[1] / fn, fa[2] / fn};
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) c->T[3 * i + j] = (i == j ? 1.0 : 0.0) - nrm[i] * nrm[j];
    double term_a = 0.0, term_b = 0.0, m[3];
    if (A->dynamic) {
        matvec3(A->rb.inertia_inv_world, ta, m);
        term_a = dot3(fa, fa) / A->rb.mass + dot3(ta, m);
    }
    if (Bd->dynamic) {
        matvec3(Bd->rb.inertia_inv_world, tb, m);
        term_b = dot3(fb, fb) / Bd->rb.mass + dot3(tb, m);
    }
    c->acc_p = dt * dt * (term_a + term_b);
    /* I/m - [c - com]x Iinv [c - com]x of each dynamic body (zero for a static one) */
    double M[2][9];
    const ref_boundary *bodies[2] = {A, Bd};
    for (int e = 0; e < 2; e++) {
        const ref_boundary *bb = bodies[e];
        if (!bb->dynamic) {
            for (int k = 0; k < 9; k++) M[e][k] = 0.0;
            continue;
        }
        double r[3], X[9], XI[9], XIX[9];
        sub3(c->center, bb->rb.com, r);
        X[0] = 0.0; X[1] = -r[2]; X[2] = r[1];
        X[3] = r[2]; X[4] = 0.0; X[5] = -r[0];
        X[6] = -r[1]; X[7] = r[0]; X[8] = 0.0;
        matmul3(X, bb->rb.inertia_inv_world, XI);
        matmul3(XI, X, XIX);
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++)
                M[e][3 * i + j] = (i == j ? 1.0 : 0.0) / bb->rb.mass - XIX[3 * i + j];
    }
    double S[9], dT[9], P[9];
    for (int k = 0; k < 9; k++) S[k] = M[0][k] + M[1][k];
    for (int k = 0; k < 9; k++) dT[k] = dt * c->T[k];
    matmul3(dT, S, P);
    c->acc_f = 1. / 3. * (((0.0 + P[0]) + P[4]) + P[8]);
}

/* Cell-level contact solve. A cell is a REF_CONTACT_REC record (body a,
 * body b, volume, center, f_a, tau_a, f_b, tau_b, a_cc pressure, a_cc friction,
 * p, lambda); a body is a REF_CONTACT_BODY record (dynamic flag, mass, com,
 * linear velocity, angular momentum, force, torque, world inverse inertia). */
#define CR_A 0
#define CR_B 1
#define CR_VOL 2
#define CR_CENTER 3
#define CR_FA 6
#define CR_TA 9
#define CR_FB 12
#define CR_TB 15
#define CR_ACCP 18
#define CR_ACCF 19
#define CR_P 20
#define CR_LAM 21
#define CB_DYN 0
#define CB_MASS 1
#define CB_COM 2
#define CB_V 5
#define CB_L 8
#define CB_F 11
#define CB_T 14
#define CB_IINV 17

static inline const double *crec(const double *cells, size_t c) {
    return cells + c * REF_CONTACT_REC;
}
static inline const double *cbody(const double *bodies, int b) {
    return bodies + (size_t)b * REF_CONTACT_BODY;
}
static inline int cell_body(const double *r, int side) { return (int)r[side ? CR_B : CR_A]; }

/* the tangential projection I - n n^T, n = f_a / max(|f_a|, 1e-15) (:927-928) */
static void contact_tangential(const double *r, double T[9]) {
    const double fn = fmax(norm3(r + CR_FA), 1e-15);
    double nrm[3] = {r[CR_FA] / fn, r[CR_FA + 1] / fn, r[CR_FA + 2] / fn};
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) T[3 * i + j] = (i == j ? 1.0 : 0.0) - nrm[i] * nrm[j];
}

/* accumulate_extra_forces (:948-986): per body over its cells in cell order */
static void contact_extra(const double *bodies, int nb, const double *cells, size_t nc,
                          const double *p, const double *lam, double *ef, double *et) {
    for (int bi = 0; bi < nb; bi++) {
        const double *own = cbody(bodies, bi);
        double f[3] = {0., 0., 0.}, t[3] = {0., 0., 0.};
        for (size_t c = 0; c < nc; c++) {
            const double *r = crec(cells, c);
            for (int side = 0; side < 2; side++) {
                if (cell_body(r, side) != bi) continue;
                const double *l = lam + 3 * c;
                const double *fs = r + (side ? CR_FB : CR_FA), *ts = r + (side ? CR_TB : CR_TA);
                double d[3], cr[3];
                for (int k = 0; k < 3; k++)
                    f[k] += side ? p[c] * fs[k] - l[k] : p[c] * fs[k] + l[k];
                if (own[CB_DYN] != 0.0) {
                    sub3(r + CR_CENTER, own + CB_COM, d);
                    cross3(d, l, cr);
                    for (int k = 0; k < 3; k++)
                        t[k] += side ? p[c] * ts[k] - cr[k] : p[c] * ts[k] + cr[k];
                }
            }
        }
        memcpy(ef + 3 * bi, f, sizeof f);
        memcpy(et + 3 * bi, t, sizeof t);
    }
}

/* compute_trial_velocities (:988-1012), RigidBodyMotion::trial_velocity and
 * trial_angular_velocity (rigid_body_motion.rs:166-181) */
static void contact_trial(const double *bodies, int nb, double dt, const double *ef,
                          const double *et, double *vt, double *wt) {
    for (int bi = 0; bi < nb; bi++) {
        const double *b = cbody(bodies, bi);
        if (b[CB_DYN] == 0.0) {
            for (int k = 0; k < 3; k++) vt[3 * bi + k] = wt[3 * bi + k] = 0.0;
            continue;
        }
        double l[3];
        for (int k = 0; k < 3; k++) {
            vt[3 * bi + k] = b[CB_V + k] + dt / b[CB_MASS] * (b[CB_F + k] + ef[3 * bi + k]);
            l[k] = b[CB_L + k] + dt * (b[CB_T + k] + et[3 * bi + k]);
        }
        matvec3(b + CB_IINV, l, wt + 3 * bi);
    }
}

/* predicted_volume_error and predicted_tangential_velocity (:996-1033) */
static void contact_predict(const double *bodies, const double *cells, size_t nc,
                            const double *T, double dt, const double *vt, const double *wt,
                            double *verr, double *vtang) {
    for (size_t c = 0; c < nc; c++) {
        const double *r = crec(cells, c);
        const int a = cell_body(r, 0), b = cell_body(r, 1);
        const double *va = vt + 3 * a, *wa = wt + 3 * a, *vb = vt + 3 * b, *wb = wt + 3 * b;
        verr[c] = -r[CR_VOL] + dt * (((dot3(va, r + CR_FA) + dot3(wa, r + CR_TA)) +
                                      dot3(vb, r + CR_FB)) +
                                     dot3(wb, r + CR_TB));
        const double *A = cbody(bodies, a), *B = cbody(bodies, b);
        if (A[CB_DYN] != 0.0 && B[CB_DYN] != 0.0) {
            double ra[3], rb[3], ca[3], cb[3], vr[3];
            sub3(r + CR_CENTER, A + CB_COM, ra);
            sub3(r + CR_CENTER, B + CB_COM, rb);
            cross3(wa, ra, ca);
            cross3(wb, rb, cb);
            for (int k = 0; k < 3; k++) vr[k] = ((va[k] + ca[k]) - vb[k]) - cb[k];
            matvec3(T + 9 * c, vr, vtang + 3 * c);
        } else {
            vtang[3 * c] = vtang[3 * c + 1] = vtang[3 * c + 2] = 0.0;
        }
    }
}

/* projected_jacobi (:1035-1102) */
static void contact_jacobi(int nb, const double *cells, size_t nc, const double *p,
                           const double *verr, const double *lam, const double *vtang,
                           uint32_t *nact, double *pn, double *ln) {
    for (int bi = 0; bi < nb; bi++) {
        uint32_t n = 0;
        for (size_t c = 0; c < nc; c++) {
            const double *r = crec(cells, c);
            int hits = (cell_body(r, 0) == bi) + (cell_body(r, 1) == bi);
            if (hits && (verr[c] < 0.0 || p[c] > 0.0)) n += (uint32_t)hits;
        }
        nact[bi] = n;
    }
    for (size_t c = 0; c < nc; c++) {
        const double *r = crec(cells, c);
        const uint32_t na = nact[cell_body(r, 0)], nb2 = nact[cell_body(r, 1)];
        const uint32_t nm = na > nb2 ? na : nb2;
        const double n_max = (double)(nm > 1u ? nm : 1u);
        const double alpha = CONTACT_RELAXATION / n_max / fmax(r[CR_ACCP], CONTACT_MIN_DIAGONAL);
        pn[c] = fmax(p[c] - alpha * verr[c], 0.0);
        const double lmax = CONTACT_FRICTION * p[c] * norm3(r + CR_FA);
        const double d = fmax(r[CR_ACCF], CONTACT_MIN_DIAGONAL);
        double raw[3], star[3], nw[3];
        for (int k = 0; k < 3; k++) raw[k] = -vtang[3 * c + k] / d;
        const double rn = norm3(raw);
        for (int k = 0; k < 3; k++) star[k] = rn > lmax ? lmax * raw[k] / rn : raw[k];
        for (int k = 0; k < 3; k++) nw[k] = lam[3 * c + k] + CONTACT_RELAXATION / n_max * star[k];
        const double nn = norm3(nw);
        for (int k = 0; k < 3; k++) ln[3 * c + k] = nn > lmax ? lmax * nw[k] / nn : nw[k];
    }
}

/* mean_fischer_burmeister (utilities/mod.rs:27-41), summed in index order */
static double contact_mean_fb(const double *a, const double *b, size_t n) {
    double sum = 0.0;
    for (size_t i = 0; i < n; i++) sum += fabs(a[i] + b[i] - sqrt(a[i] * a[i] + b[i] * b[i]));
    return sum / (double)n;
}

/* the three merit inputs of solve_for_p_c's termination test (:310-334) */
static void contact_merit(const double *cells, size_t nc, const double *p, const double *lam,
                          const double *vtang, double *pp, double *vtn, double *dev) {
    for (size_t c = 0; c < nc; c++) {
        const double *r = crec(cells, c);
        pp[c] = p[c] * fmax(r[CR_ACCP], CONTACT_MIN_DIAGONAL);
        vtn[c] = norm3(vtang + 3 * c);
        dev[c] = fmax(r[CR_ACCF], CONTACT_MIN_DIAGONAL) *
                 (CONTACT_FRICTION * p[c] * norm3(r + CR_FA) - norm3(lam + 3 * c));
    }
}

/* solve_for_p_c (:186-637) after compute_diagonal: NNCG-accelerated projected
 * Jacobi over the cells; writes p and lambda into the cell records and the
 * committed per-body force and torque into extra[6 * nb]; returns the
 * iteration count */
uint32_t ref_contact_solve(const double *bodies, int nb, double *cells, size_t nc, double dt,
                           double *extra) {
    for (size_t k = 0; k < 6 * (size_t)nb; k++) extra[k] = 0.0;
    if (nc == 0) return 0;
    double *T = (double *)malloc(9 * nc * sizeof(double));
    for (size_t c = 0; c < nc; c++) contact_tangential(crec(cells, c), T + 9 * c);
    double *p = (double *)calloc(nc, sizeof(double)), *verr = (double *)calloc(nc, sizeof(double));
    double *lam = (double *)calloc(3 * nc, sizeof(double)), *vtang = (double *)calloc(3 * nc, sizeof(double));
    double *sp = (double *)calloc(nc, sizeof(double)), *sl = (double *)calloc(3 * nc, sizeof(double));
    double *pj = (double *)calloc(nc, sizeof(double)), *lj = (double *)calloc(3 * nc, sizeof(double));
    double *gp = (double *)calloc(nc, sizeof(double)), *gl = (double *)calloc(3 * nc, sizeof(double));
    double *pp = (double *)calloc(nc, sizeof(double)), *vtn = (double *)calloc(nc, sizeof(double));
    double *dev = (double *)calloc(nc, sizeof(double));
    double *ef = (double *)calloc(3 * (size_t)nb, sizeof(double)), *et = (double *)calloc(3 * (size_t)nb, sizeof(double));
    double *vt = (double *)calloc(3 * (size_t)nb, sizeof(double)), *wt = (double *)calloc(3 * (size_t)nb, sizeof(double));
    uint32_t *nact = (uint32_t *)calloc((size_t)nb, sizeof(uint32_t));

    contact_extra(bodies, nb, cells, nc, p, lam, ef, et);
    contact_trial(bodies, nb, dt, ef, et, vt, wt);
    contact_predict(bodies, cells, nc, T, dt, vt, wt, verr, vtang);
    contact_jacobi(nb, cells, nc, p, verr, lam, vtang, nact, pj, lj);
    double prev_p2 = 0.0, prev_l2 = 0.0;
    for (size_t c = 0; c < nc; c++) {
        gp[c] = -(pj[c] - p[c]);
        for (int k = 0; k < 3; k++) gl[3 * c + k] = -(lj[3 * c + k] - lam[3 * c + k]);
    }
    for (size_t c = 0; c < nc; c++) {
        sp[c] = -gp[c];
        for (int k = 0; k < 3; k++) sl[3 * c + k] = -gl[3 * c + k];
    }
    memcpy(p, pj, nc * sizeof(double));
    memcpy(lam, lj, 3 * nc * sizeof(double));
    for (size_t c = 0; c < nc; c++) prev_p2 += gp[c] * gp[c];
    for (size_t c = 0; c < nc; c++) prev_l2 += dot3(gl + 3 * c, gl + 3 * c);
    contact_extra(bodies, nb, cells, nc, p, lam, ef, et);
    contact_trial(bodies, nb, dt, ef, et, vt, wt);
    contact_predict(bodies, cells, nc, T, dt, vt, wt, verr, vtang);
    contact_merit(cells, nc, p, lam, vtang, pp, vtn, dev);

    uint32_t it = 0;
    while (!(contact_mean_fb(pp, verr, nc) < CONTACT_FB_TOL_PRESSURE &&
             contact_mean_fb(vtn, dev, nc) < CONTACT_FB_TOL_FRICTION)) {
        if (it >= CONTACT_MAX_ITERATIONS) break;
        contact_jacobi(nb, cells, nc, p, verr, lam, vtang, nact, pj, lj);
        double gp2 = 0.0, gl2 = 0.0;
        for (size_t c = 0; c < nc; c++) {
            gp[c] = -(pj[c] - p[c]);
            for (int k = 0; k < 3; k++) gl[3 * c + k] = -(lj[3 * c + k] - lam[3 * c + k]);
        }
        for (size_t c = 0; c < nc; c++) gp2 += gp[c] * gp[c];
        for (size_t c = 0; c < nc; c++) gl2 += dot3(gl + 3 * c, gl + 3 * c);
        const double beta = (gp2 + gl2) / fmax(prev_p2 + prev_l2, 2.2250738585072014e-308);
        if (beta > 1.0) { /* NNCG restart */
            for (size_t c = 0; c < nc; c++) {
                sp[c] = 0.0;
                for (int k = 0; k < 3; k++) sl[3 * c + k] = 0.0;
            }
            memcpy(p, pj, nc * sizeof(double));
            memcpy(lam, lj, 3 * nc * sizeof(double));
        } else {
            for (size_t c = 0; c < nc; c++) {
                /* friction first: its cone bound uses the pressure before this update */
                double ln[3], tp[3];
                for (int k = 0; k < 3; k++) ln[k] = lj[3 * c + k] + beta * sl[3 * c + k];
                matvec3(T + 9 * c, ln, tp);
                const double mag = norm3(tp);
                const double lmax = CONTACT_FRICTION * p[c] * norm3(crec(cells, c) + CR_FA);
                if (mag > lmax) {
                    for (int k = 0; k < 3; k++) {
                        lam[3 * c + k] = lj[3 * c + k];
                        sl[3 * c + k] = 0.0;
                    }
                } else {
                    for (int k = 0; k < 3; k++) {
                        lam[3 * c + k] = tp[k];
                        sl[3 * c + k] = beta * sl[3 * c + k] - gl[3 * c + k];
                    }
                }
            }
            for (size_t c = 0; c < nc; c++) {
                const double pn = pj[c] + beta * sp[c];
                if (pn < 0.0) {
                    p[c] = pj[c];
                    sp[c] = 0.0;
                } else {
                    p[c] = pn;
                    sp[c] = beta * sp[c] - gp[c];
                }
            }
        }
        prev_p2 = gp2;
        prev_l2 = gl2;
        contact_extra(bodies, nb, cells, nc, p, lam, ef, et);
        contact_trial(bodies, nb, dt, ef, et, vt, wt);
        contact_predict(bodies, cells, nc, T, dt, vt, wt, verr, vtang);
        contact_merit(cells, nc, p, lam, vtang, pp, vtn, dev);
        it++;
    }
    contact_extra(bodies, nb, cells, nc, p, lam, ef, et);
    for (int bi = 0; bi < nb; bi++)
        for (int k = 0; k < 3; k++) {
            extra[6 * bi + k] = ef[3 * bi + k];
            extra[6 * bi + 3 + k] = et[3 * bi + k];
        }
    for (size_t c = 0; c < nc; c++) {
        double *r = cells + c * REF_CONTACT_REC;
        r[CR_P] = p[c];
        for (int k = 0; k < 3; k++) r[CR_LAM + k] = lam[3 * c + k];
    }
    free(T); free(p); free(verr); free(lam); free(vtang); free(sp); free(sl); free(pj); free(lj);
    free(gp); free(gl); free(pp); free(vtn); free(dev); free(ef); free(et); free(vt); free(wt);
    free(nact);
    return it;
}

/* solve_for_p_c of a system's active cells: compute_diagonal per cell, the
 * cell-level solve, then add_force_and_torque on every dynamic body
 * (rigid_body_motion.rs:156-159) */
static void contact_solve(ref_system *s, contact_cell *cells, size_t nc, double dt) {
    s->contact_iters = 0;
    if (nc == 0) return;
    const int nb = s->nb;
    for (size_t c = 0; c < nc; c++) contact_diagonal(s, &cells[c], dt);
    double *bodies = (double *)calloc((size_t)nb * REF_CONTACT_BODY, sizeof(double));
    for (int bi = 0; bi < nb; bi++) {
        const ref_boundary *b = &s->b[bi];
        double *o = bodies + (size_t)bi * REF_CONTACT_BODY;
        o[CB_DYN] = b->dynamic ? 1.0 : 0.0;
        if (!b->dynamic) continue;
        o[CB_MASS] = b->rb.mass;
        for (int k = 0; k < 3; k++) {
            o[CB_COM + k] = b->rb.com[k];
            o[CB_V + k] = b->rb.lin_vel[k];
            o[CB_L + k] = b->rb.ang_mom[k];
            o[CB_F + k] = b->rb.force[k];
            o[CB_T + k] = b->rb.torque[k];
        }
        memcpy(o + CB_IINV, b->rb.inertia_inv_world, 9 * sizeof(double));
    }
    s->contact_rec = (double *)realloc(s->contact_rec, nc * REF_CONTACT_REC * sizeof(double));
    for (size_t c = 0; c < nc; c++) {
        const contact_cell *cc = &cells[c];
        double *r = s->contact_rec + c * REF_CONTACT_REC;
        r[CR_A] = (double)cc->a;
        r[CR_B] = (double)cc->b;
        r[CR_VOL] = cc->volume;
        for (int k = 0; k < 3; k++) {
            r[CR_CENTER + k] = cc->center[k];
            r[CR_FA + k] = cc->fa[k];
            r[CR_TA + k] = cc->taua[k];
            r[CR_FB + k] = cc->fb[k];
            r[CR_TB + k] = cc->taub[k];
            r[CR_LAM + k] = 0.0;
        }
        r[CR_ACCP] = cc->acc_p;
        r[CR_ACCF] = cc->acc_f;
        r[CR_P] = 0.0;
    }
    double *extra = (double *)calloc(6 * (size_t)nb, sizeof(double));
    s->contact_iters = ref_contact_solve(bodies, nb, s->contact_rec, nc, dt, extra);
    for (int bi = 0; bi < nb; bi++) {
        ref_boundary *b = &s->b[bi];
        if (!b->dynamic) continue;
        for (int k = 0; k < 3; k++) {
            b->rb.force[k] += extra[6 * bi + k];
            b->rb.torque[k] += extra[6 * bi + 3 + k];
        }
    }
    free(bodies);
    free(extra);
}

/* contact handling of a system whose boundaries are all volume maps */
static void contact_step(ref_system *s, double dt) {
    s->contact_cells = 0;
    s->contact_iters = 0;
    if (s->contact_degree <= 0 || s->nb < 2) return;
    int dyn = 0;
    for (int i = 0; i < s->nb; i++) {
        if (s->b[i].kind != 1) return;
        dyn |= s->b[i].dynamic;
    }
    if (!dyn) return;
    size_t nc = 0;
    contact_cell *cells = contact_generate(s, &nc);
    s->contact_cells = (uint32_t)nc;
    contact_solve(s, cells, nc, dt);
    for (size_t c = 0; c < nc; c++) free(cells[c].qp);
    free(cells);
}

int ref_system_set_aabb(ref_system *s, int b, const double center[3], const double half[3]) {
    if (b < 0 || b >= s->nb) return -1;
    s->b[b].has_aabb = 1;
    memcpy(s->b[b].aabb_c, center, 3 * sizeof(double));
    memcpy(s->b[b].aabb_h, half, 3 * sizeof(double));
    return 0;
}

int ref_system_set_contact(ref_system *s, const double *gl_x, const double *gl_w, int degree) {
    if (degree < 0 || degree > 32 || s->nb > 64) return -1;
    s->contact_degree = degree;
    for (int k = 0; k < degree; k++) {
        s->contact_x[k] = gl_x[k];
        s->contact_w[k] = gl_w[k];
    }
    return 0;
}

void ref_system_contact_stats(const ref_system *s, uint32_t out[2]) {
    out[0] = s->contact_cells;
    out[1] = s->contact_iters;
}

size_t ref_system_contact_cells(const ref_system *s, double *out, size_t cap) {
    size_t n = s->contact_cells;
    if (out) memcpy(out, s->contact_rec, (n < cap ? n : cap) * REF_CONTACT_REC * sizeof(double));
    return n;
}

/* IntegrationScheme::integrate (sph/integration_schemes) */
static void integrate(ref_system *s, double dt) {
    if (s->P.integrator == REF_EULER_CROMER) { /* euler_cromer.rs:12-25 */
        for (size_t i = 0; i < 3 * s->n; i++) {
            s->v[i] = s->v[i] + dt * s->a[i];
            s->x[i] = s->x[i] + dt * s->v[i];
        }
    } else if (s->P.integrator == REF_EXPLICIT_EULER) { /* explicit_euler.rs:12-25 */
        for (size_t i = 0; i < 3 * s->n; i++) {
            const double x0 = s->x[i], v0 = s->v[i];
            s->x[i] = x0 + dt * v0;
            s->v[i] = v0 + dt * s->a[i];
        }
    } else if (s->P.integrator == REF_VERLET) { /* verlet.rs:13-34; slot = x(t - dt) */
        for (size_t i = 0; i < 3 * s->n; i++) {
            const double x0 = s->x[i];
            const double nx = x0 + (x0 - s->ix[i]) + powi2(dt) * s->a[i];
            s->v[i] = (nx - x0) / dt;
            s->ix[i] = x0;
            s->x[i] = nx;
        }
    } else { /* take_predicted.rs:14-27 */
        double *t;
        t = s->x; s->x = s->ix; s->ix = t;
        t = s->v; s->v = s->iv; s->iv = t;
    }
}

/* SPHSystem::step_forward_in_time   (sph/mod.rs:125-144) */
int ref_system_step(ref_system *s, int nsteps) {
    for (int st = 0; st < nsteps; st++) {
        double dt = s->dt;
        integrate(s, dt);
        contact_step(s, dt); /* VolumeMapBoundary::step_forward_in_time (:1440-1453) */
        for (int i = 0; i < s->nb; i++) { /* B::step_forward_in_time */
            ref_boundary *b = &s->b[i];
            if (!b->dynamic) continue;
            ref_rigid_step(&b->rb, dt);
            if (b->kind == 0) ss_update_positions(b);
        }
        s->steps++;
        if (s->P.cfl) /* cfl_time_step: current_time += dt (sph/mod.rs:134-139) */
            s->time += dt;
        else
            s->time = s->time_offset + (double)(s->steps - s->offset_steps) * s->P.time_increment;
        update(s);
    }
    return 0;
}

size_t ref_system_len(const ref_system *s) { return s->n; }

void ref_system_get(const ref_system *s, double *x, double *v, double *a, double *V,
                    double *p, double *m, uint32_t *fid, uint64_t *origin) {
    size_t n = s->n;
    if (x) memcpy(x, s->x, 3 * n * sizeof(double));
    if (v) memcpy(v, s->v, 3 * n * sizeof(double));
    if (a) memcpy(a, s->a, 3 * n * sizeof(double));
    if (V) memcpy(V, s->V, n * sizeof(double));
    if (p) memcpy(p, s->p, n * sizeof(double));
    if (m) memcpy(m, s->m, n * sizeof(double));
    if (fid) memcpy(fid, s->fid, n * sizeof(uint32_t));
    if (origin) memcpy(origin, s->origin, n * sizeof(uint64_t));
}

/* take_measurement   (sph/mod.rs:147-184, reductions :489-564) */
void ref_system_measure(const ref_system *s, ref_measurement *o) {
    memset(o, 0, sizeof *o);
    o->time = s->time;
    o->time_step_size = s->dt;
    o->density = s->avg_density;
    double se = 0., ke = 0.;
    for (size_t i = 0; i < s->n; i++) {
        se += (s->V[i] < s->rest_volume) ? s->rest_volume / s->V[i] : 1.;
        ke += 0.5 * s->m[i] * norm_sq3(s->v + 3 * i);
    }
    o->density_error = s->n ? 100. * (se / (double)s->n - 1.) : 0.0;
    o->kinetic_energy = s->n ? ke / (double)s->n : 0.0;
    const int state_equation = s->P.solver == REF_SESPH || s->P.solver == REF_SESPH_SPLITTING;
    o->solver_iterations = state_equation ? 0 : s->last_iter;
    o->eqs1_iterations = s->P.solver == REF_IISPHWOST ? s->eqs1_iter : 0;
    o->predicted_density_error = state_equation ? 0.0 : s->last_err;
    o->steps = s->steps;
    o->num_particles = s->n;
}

size_t ref_system_error_history(const ref_system *s, double *out, size_t cap) {
    size_t k = s->err_hist_len < cap ? s->err_hist_len : cap;
    memcpy(out, s->err_hist, k * sizeof(double));
    return s->err_hist_len;
}

int ref_system_body_state(const ref_system *s, int bi, double out[22]) {
    if (bi < 0 || bi >= s->nb) return -1;
    const ref_rigid *rb = &s->b[bi].rb;
    memcpy(out, rb->com, 3 * sizeof(double));
    memcpy(out + 3, rb->q, 4 * sizeof(double));
    memcpy(out + 7, rb->lin_vel, 3 * sizeof(double));
    memcpy(out + 10, rb->ang_mom, 3 * sizeof(double));
    memcpy(out + 13, rb->force, 3 * sizeof(double));
    memcpy(out + 16, rb->torque, 3 * sizeof(double));
    memcpy(out + 19, rb->ang_vel, 3 * sizeof(double));
    return 0;
}

int ref_system_body_forces(const ref_system *s, int bi, double out[6]) {
    if (bi < 0 || bi >= s->nb) return -1;
    memcpy(out, s->b[bi].applied, 6 * sizeof(double));
    return 0;
}

const ref_nlist *ref_system_fluid_nlist(const ref_system *s) { return &s->nl; }

size_t ref_system_boundary_samples(const ref_system *s, int bi, size_t i, int mode,
                                   double *out, size_t cap) {
    const ref_boundary *b = &s->b[bi];
    uint64_t lo = b->off[mode][i], hi = b->off[mode][i + 1];
    size_t k = 0;
    for (uint64_t e = lo; e < hi; e++, k++) {
        if (k >= cap) continue;
        uint64_t id = b->ids[mode][e];
        memcpy(out + 7 * k, b->pos + 3 * id, 3 * sizeof(double));
        memcpy(out + 7 * k + 3, b->vel + 3 * id, 3 * sizeof(double));
        out[7 * k + 6] = b->vol[id];
    }
    return (size_t)(hi - lo);
}

/* SPHSystem::get_quantity_at_positions (sph/mod.rs:241-341) with the
 * quantities of quantities.rs: get_speed (:64-120), get_volume (:14-61),
 * get_density (:123-170), get_density_error (:173-228; the caller passes the
 * fluid masses as its `sample_volumes`, sph/mod.rs:297-306, restated as is),
 * get_pressure (:231-280), get_kinetic_energy (:283-330). Queries are f32
 * positions widened to f64; results are narrowed to f32. The boundaries'
 * sample lists are rebuilt for the queries and then for the fluid again. */
int ref_system_quantity_at_positions(ref_system *s, int quantity, size_t nq, const float *pos,
                                     float *out) {
    if (quantity < 0 || quantity > 5) return -1;
    const double h = s->P.kernel_support_radius, V0 = s->rest_volume;
    double *q = dalloc(3 * nq);
    for (size_t k = 0; k < 3 * nq; k++) q[k] = (double)pos[k];
    ref_nlist nl;
    memset(&nl, 0, sizeof nl);
    ref_find_samples(h, h, q, nq, s->x, s->n, &nl);
    double *x0 = s->x;
    const size_t n0 = s->n;
    s->x = q;
    s->n = nq;
    find_boundary_samples(s);
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < nq; i++) {
        const double *pi = q + 3 * i;
        double res = 0.0;
        if (quantity == 0) { /* Speed */
            double acc[3] = {0.0, 0.0, 0.0};
            for (uint64_t e = nl.offsets[i]; e < nl.offsets[i + 1]; e++) {
                const uint64_t j = nl.indices[e];
                double r[3];
                sub3(pi, x0 + 3 * j, r);
                const double w = ref_kernel_function(r, h);
                for (int k = 0; k < 3; k++) acc[k] += (s->v[3 * j + k] * s->V[j]) * w;
            }
            for (int bi = 0; bi < s->nb; bi++) {
                const ref_boundary *b = &s->b[bi];
                for (uint64_t e = b->off[0][i]; e < b->off[0][i + 1]; e++) {
                    const uint64_t id = b->ids[0][e];
                    double r[3];
                    sub3(pi, b->pos + 3 * id, r);
                    const double w = ref_kernel_function(r, h);
                    for (int k = 0; k < 3; k++) acc[k] += (b->vel[3 * id + k] * b->vol[id]) * w;
                }
            }
            res = norm3(acc);
        } else if (quantity == 1) { /* Volume */
            double acc = 0.0;
            for (uint64_t e = nl.offsets[i]; e < nl.offsets[i + 1]; e++) {
                double r[3];
                sub3(x0 + 3 * nl.indices[e], pi, r);
                acc += V0 * ref_kernel_function(r, h);
            }
            for (int bi = 0; bi < s->nb; bi++) {
                const ref_boundary *b = &s->b[bi];
                for (uint64_t e = b->off[0][i]; e < b->off[0][i + 1]; e++) {
                    const uint64_t id = b->ids[0][e];
                    double r[3];
                    sub3(b->pos + 3 * id, pi, r);
                    acc += b->vol[id] * ref_kernel_function(r, h);
                }
            }
            res = V0 / acc;
        } else {
            double acc = 0.0;
            for (uint64_t e = nl.offsets[i]; e < nl.offsets[i + 1]; e++) {
                const uint64_t j = nl.indices[e];
                double r[3];
                sub3(pi, x0 + 3 * j, r);
                const double w = ref_kernel_function(r, h);
                if (quantity == 2) { /* Density */
                    acc += s->m[j] * w;
                } else if (quantity == 3) { /* DensityError, masses as volumes */
                    if (!(s->m[j] < V0)) continue;
                    const double err = V0 / s->m[j] - 1.;
                    acc += (err * s->m[j]) * w;
                } else if (quantity == 4) { /* Pressure */
                    acc += (s->p[j] * s->V[j]) * w;
                } else { /* KineticEnergy */
                    acc += (((0.5 * s->m[j]) * norm_sq3(s->v + 3 * j)) * s->V[j]) * w;
                }
            }
            res = quantity == 3 ? 100. * acc : acc;
        }
        out[i] = (float)res;
    }
    s->x = x0;
    s->n = n0;
    find_boundary_samples(s);
    ref_nlist_free(&nl);
    free(q);
    return 0;
}

double ref_system_static_volume(const ref_system *s, int bi, size_t k) {
    return s->b[bi].svol[k];
}

/* ===================================================================== */
/* test hooks: one reference function at a time on the current state, so  */
/* the reference's own per-function tests (given volumes, pressures,      */
/* velocities) can be restated against the oracle                          */
/* ===================================================================== */
int ref_system_call(ref_system *s, int fn) {
    switch (fn) {
        case REF_FN_NEIGHBORS: { /* the search part of update() (sph/mod.rs:701-712) */
            double h = s->P.kernel_support_radius;
            ref_nlist_free(&s->nl);
            ref_find_samples(h, h, s->x, s->n, s->x, s->n, &s->nl);
            for (int i = 0; i < s->nb; i++) memset(s->b[i].applied, 0, sizeof s->b[i].applied);
            find_boundary_samples(s);
            return 0;
        }
        case REF_FN_VOLUME: get_volume(s); return 0;
        case REF_FN_RESET_ACCELERATION: memset(s->a, 0, 3 * s->n * sizeof(double)); return 0;
        case REF_FN_GRAVITY: add_gravity(s); return 0;
        case REF_FN_VISCOSITY: add_viscosity(s); return 0;
        case REF_FN_PRESSURE_ACCELERATION: add_pressure_acceleration(s, s->a, 0, 1); return 0;
        case REF_FN_PRESSURE_ACCELERATION_OVERWRITE:
            add_pressure_acceleration(s, s->a, 1, 1);
            return 0;
        case REF_FN_PRESSURE_ACCELERATION_SCRATCH:
            add_pressure_acceleration(s, s->pacc, 1, 0);
            return 0;
        case REF_FN_SESPH: sesph_solve(s); return 0;
        case REF_FN_PRED_VEL: set_pred_vel(s, 0); return 0;
        case REF_FN_SOURCE_VDE: set_source_term(s, 0); return 0;
        case REF_FN_SOURCE_VP: set_source_term(s, 1); return 0;
        case REF_FN_DIAGONAL: set_diagonal(s); return 0;
        case REF_FN_INITIALIZE: iisph_initialize(s); return 0;
        case REF_FN_RESOLVE: return (int)resolve_pressure(s, -1);
        case REF_FN_SESPH_SPLITTING: sesph_split_solve(s); return 0;
        case REF_FN_INTEGRATE: integrate(s, s->dt); return 0;
        default: return -1;
    }
}

int ref_system_field(ref_system *s, int field, int write, double *buf) {
    double *f = NULL;
    size_t k = 0;
    switch (field) {
        case REF_F_X: f = s->x; k = 3 * s->n; break;
        case REF_F_VEL: f = s->v; k = 3 * s->n; break;
        case REF_F_ACC: f = s->a; k = 3 * s->n; break;
        case REF_F_MASS: f = s->m; k = s->n; break;
        case REF_F_VOLUME: f = s->V; k = s->n; break;
        case REF_F_PRESSURE: f = s->p; k = s->n; break;
        case REF_F_PRED_VEL: f = s->sv; k = 3 * s->n; break;
        case REF_F_PRED_POS: f = s->sx; k = 3 * s->n; break;
        case REF_F_SOURCE: f = s->s_f; k = s->n; break;
        case REF_F_DIAGONAL: f = s->a_ff; k = s->n; break;
        case REF_F_PRESSURE_ACC: f = s->pacc; k = 3 * s->n; break;
        case REF_F_TIME_STEP: f = &s->dt; k = 1; break;
        case REF_F_SLOT_POS: f = s->ix; k = 3 * s->n; break;
        default: return -1;
    }
    if (write)
        memcpy(f, buf, k * sizeof(double));
    else
        memcpy(buf, f, k * sizeof(double));
    return (int)k;
}

int ref_system_sample_volumes(ref_system *s, int bi, int write, double *buf) {
    if (bi < 0 || bi >= s->nb || s->b[bi].kind != 0) return -1;
    ref_boundary *b = &s->b[bi];
    if (write)
        memcpy(b->svol, buf, b->ns * sizeof(double));
    else
        memcpy(buf, b->svol, b->ns * sizeof(double));
    return (int)b->ns;
}

void ref_system_set_threads(int nthreads) {
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
}


/* =========================================================================
 * Volume-map construction (setup of VolumeMapBoundary)
 * ========================================================================= */

/* cubic_extension_fn (volume_map_boundary.rs:1716-1726):
 * W((sd, 0, 0), h) / W(0, h) on (0, h), 1 below, 0 above */
static double ref_cubic_extension(double sd, double h) {
    if (sd >= h) return 0.;
    if (sd <= 0.) return 1.;
    double r[3] = {sd, 0., 0.}, z[3] = {0., 0., 0.};
    return ref_kernel_function(r, h) / ref_kernel_function(z, h);
}

/* ball_volume (volume_map_boundary.rs:34-36) */
static double ref_ball_volume(double h) {
    return 4.0 / 3.0 * 3.14159265358979323846 * ((h * h) * h);
}

int ref_vm_integrate_ball(const ref_lattice *sdf, const double center[3], double h,
                          const double *gl_x, const double *gl_w, int degree, double *out) {
    const double PI = 3.14159265358979323846;
    const int n = degree;
    double acc_r = 0.0;
    for (int i = 0; i < n; i++) {
        const double r = 0.5 * ((h - 0.0) * gl_x[i] + (h + 0.0));
        double acc_t = 0.0;
        for (int j = 0; j < n; j++) {
            const double theta = 0.5 * ((PI - 0.0) * gl_x[j] + (PI + 0.0));
            const double st = sin(theta), ct = cos(theta);
            double acc_p = 0.0;
            for (int k = 0; k < n; k++) {
                const double phi = 0.5 * ((2. * PI - 0.0) * gl_x[k] + (2. * PI + 0.0));
                const double p[3] = {center[0] + r * st * cos(phi), center[1] + r * st * sin(phi),
                                     center[2] + r * ct};
                double sd;
                int e = ref_serendipity_function(sdf, p, &sd);
                if (e) return e;
                acc_p += ref_cubic_extension(sd, h) * gl_w[k];
            }
            const double inner_p = 0.5 * (2. * PI - 0.0) * acc_p;
            acc_t += st * inner_p * gl_w[j];
        }
        const double inner_t = 0.5 * (PI - 0.0) * acc_t;
        acc_r += (r * r) * inner_t * gl_w[i];
    }
    *out = 0.5 * (h - 0.0) * acc_r;
    return 0;
}

int ref_vm_node_value(const ref_lattice *sdf, const double p[3], double h, const double *gl_x,
                      const double *gl_w, int degree, double *out) {
    const double bv = ref_ball_volume(h);
    double sd, v;
    if (ref_serendipity_function(sdf, p, &sd)) {
        v = 0.0;  /* pruned cell or outside -> treated as outside the band */
    } else if (sd >= 2.0 * h) {
        v = 0.0;
    } else if (sd <= -h) {
        v = 4.0 / 3.0 * 3.14159265358979323846 * ((h * h) * h);
    } else {
        int e = ref_vm_integrate_ball(sdf, p, h, gl_x, gl_w, degree, &v);
        if (e) return e;
    }
    *out = v < 0.0 ? 0.0 : (v > bv ? bv : v);  /* f64::clamp(0, bv) */
    return 0;
}

uint64_t ref_vm_build(const ref_lattice *sdf, double h, const double x_min[3], const int64_t n[3],
                  double dx, const double *gl_x, const double *gl_w, int degree, double *values,
                  uint8_t *present) {
    const int64_t nx = 3 * n[0] + 1, ny = 3 * n[1] + 1, nz = 3 * n[2] + 1;
    const size_t nodes = (size_t)(nx * ny * nz);
    uint8_t *ok = (uint8_t *)calloc(nodes, 1);
    uint8_t *isnode = (uint8_t *)calloc(nodes, 1);
    for (int64_t cz = 0; cz < n[2]; cz++)
        for (int64_t cy = 0; cy < n[1]; cy++)
            for (int64_t cx = 0; cx < n[0]; cx++)
                for (int i = 0; i < 32; i++)
                    isnode[((3 * cz + NODE_OFFS[i][2]) * ny + (3 * cy + NODE_OFFS[i][1])) * nx +
                           (3 * cx + NODE_OFFS[i][0])] = 1;
    uint64_t quadratures = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : quadratures)
    for (int64_t k = 0; k < (int64_t)nodes; k++) {
        values[k] = 0.0;
        present[k] = 0;
        if (!isnode[k]) continue;
        const int64_t ix = k % nx, iy = (k / nx) % ny, iz = k / (nx * ny);
        /* x_min + key / 3 * dx (discretization.rs:171-175) */
        const double p[3] = {x_min[0] + (double)ix / 3.0 * dx, x_min[1] + (double)iy / 3.0 * dx,
                             x_min[2] + (double)iz / 3.0 * dx};
        double v, sc;
        if (ref_serendipity_function(sdf, p, &sc) == 0 && sc < 2.0 * h && sc > -h) quadratures++;
        if (ref_vm_node_value(sdf, p, h, gl_x, gl_w, degree, &v) == 0) {
            values[k] = v;
            ok[k] = 1;
        }
    }
    /* a cell is kept iff all 32 nodes evaluated (values are >= 0, so the
     * lower bound 0 prunes nothing else) */
    for (int64_t cz = 0; cz < n[2]; cz++)
        for (int64_t cy = 0; cy < n[1]; cy++)
            for (int64_t cx = 0; cx < n[0]; cx++) {
                int good = 1;
                for (int i = 0; i < 32 && good; i++)
                    good = ok[((3 * cz + NODE_OFFS[i][2]) * ny + (3 * cy + NODE_OFFS[i][1])) * nx +
                              (3 * cx + NODE_OFFS[i][0])];
                if (!good) continue;
                for (int i = 0; i < 32; i++)
                    present[((3 * cz + NODE_OFFS[i][2]) * ny + (3 * cy + NODE_OFFS[i][1])) * nx +
                            (3 * cx + NODE_OFFS[i][0])] = 1;
            }
    for (size_t k = 0; k < nodes; k++)
        if (!present[k]) values[k] = 0.0;
    free(ok);
    free(isnode);
    return quadratures;
}

/* ===================================================================== */
/* signed distance to an oriented triangle mesh                            */
/* (TriangleMeshWrapper::signed_distance, volume_map_boundary.rs:1655-1671) */
/* ===================================================================== */
/* The reference projects onto parry3d's oriented TriMesh (closest point and
 * feature) and signs the distance with the feature's outward pseudo-normal.
 * Restated with published methods: the closest point of each triangle by
 * Voronoi-region tests (Ericson, Real-Time Collision Detection, 5.1.5), the
 * first triangle of least distance in index order, and the pseudo-normals of
 * Baerentzen and Aanaes (2005): the face normal, the normalized sum of the
 * two face normals at an edge, and the angle-weighted normalized sum of the
 * face normals at a vertex. */
struct ref_mesh {
    size_t nv, nt, ne;
    double *v;   /* [nv][3] */
    uint32_t *t; /* [nt][3] */
    double *fn;  /* [nt][3] unit face normals */
    double *vn;  /* [nv][3] vertex pseudo-normals */
    double *en;  /* [ne][3] edge pseudo-normals */
    uint32_t *te; /* [nt][3] edge ids of (v0 v1), (v1 v2), (v2 v0) */
};

static void normalize3(double *a) {
    const double l = norm3(a);
    if (l > 0.0)
        for (int k = 0; k < 3; k++) a[k] = a[k] / l;
}

typedef struct mesh_edge_ref {
    uint32_t lo, hi, tri, local;
} mesh_edge_ref;

static int mesh_edge_cmp(const void *x, const void *y) {
    const mesh_edge_ref *a = (const mesh_edge_ref *)x, *b = (const mesh_edge_ref *)y;
    if (a->lo != b->lo) return a->lo < b->lo ? -1 : 1;
    if (a->hi != b->hi) return a->hi < b->hi ? -1 : 1;
    if (a->tri != b->tri) return a->tri < b->tri ? -1 : 1;
    return a->local < b->local ? -1 : (a->local > b->local);
}

ref_mesh *ref_mesh_create(const double *v, size_t nv, const uint32_t *t, size_t nt) {
    for (size_t i = 0; i < 3 * nt; i++)
        if (t[i] >= nv) return NULL;
    ref_mesh *m = (ref_mesh *)calloc(1, sizeof(ref_mesh));
    m->nv = nv;
    m->nt = nt;
    m->v = (double *)malloc(3 * nv * sizeof(double) + 1);
    m->t = (uint32_t *)malloc(3 * nt * sizeof(uint32_t) + 1);
    memcpy(m->v, v, 3 * nv * sizeof(double));
    memcpy(m->t, t, 3 * nt * sizeof(uint32_t));
    m->fn = (double *)calloc(3 * nt + 1, sizeof(double));
    m->vn = (double *)calloc(3 * nv + 1, sizeof(double));
    m->te = (uint32_t *)calloc(3 * nt + 1, sizeof(uint32_t));
    for (size_t f = 0; f < nt; f++) {
        const double *a = v + 3 * t[3 * f], *b = v + 3 * t[3 * f + 1], *c = v + 3 * t[3 * f + 2];
        double e1[3], e2[3];
        sub3(b, a, e1);
        sub3(c, a, e2);
        cross3(e1, e2, m->fn + 3 * f);
        normalize3(m->fn + 3 * f);
    }
    /* vertex pseudo-normals: incident faces in index order, weighted by the
     * face's angle at the vertex */
    for (size_t f = 0; f < nt; f++)
        for (int i = 0; i < 3; i++) {
            const double *p = v + 3 * t[3 * f + i], *q = v + 3 * t[3 * f + (i + 1) % 3],
                         *r = v + 3 * t[3 * f + (i + 2) % 3];
            double e1[3], e2[3];
            sub3(q, p, e1);
            sub3(r, p, e2);
            const double l = norm3(e1) * norm3(e2);
            double cs = l > 0.0 ? dot3(e1, e2) / l : 1.0;
            cs = cs < -1.0 ? -1.0 : (cs > 1.0 ? 1.0 : cs);
            const double ang = acos(cs);
            double *o = m->vn + 3 * t[3 * f + i];
            for (int k = 0; k < 3; k++) o[k] += ang * m->fn[3 * f + k];
        }
    for (size_t i = 0; i < nv; i++) normalize3(m->vn + 3 * i);
    /* edges: unique vertex pairs; the sum of the incident face normals in
     * face order */
    mesh_edge_ref *er = (mesh_edge_ref *)malloc(3 * nt * sizeof(mesh_edge_ref) + 1);
    for (size_t f = 0; f < nt; f++)
        for (uint32_t i = 0; i < 3; i++) {
            const uint32_t a = t[3 * f + i], b = t[3 * f + (i + 1) % 3];
            mesh_edge_ref e = {a < b ? a : b, a < b ? b : a, (uint32_t)f, i};
            er[3 * f + i] = e;
        }
    qsort(er, 3 * nt, sizeof(mesh_edge_ref), mesh_edge_cmp);
    m->en = (double *)calloc(3 * 3 * nt + 1, sizeof(double));
    size_t ne = 0;
    for (size_t k = 0; k < 3 * nt; k++) {
        if (k > 0 && (er[k].lo != er[k - 1].lo || er[k].hi != er[k - 1].hi)) ne++;
        m->te[3 * er[k].tri + er[k].local] = (uint32_t)ne;
        for (int d = 0; d < 3; d++) m->en[3 * ne + d] += m->fn[3 * er[k].tri + d];
    }
    m->ne = nt ? ne + 1 : 0;
    for (size_t e = 0; e < m->ne; e++) normalize3(m->en + 3 * e);
    free(er);
    return m;
}

void ref_mesh_destroy(ref_mesh *m) {
    if (!m) return;
    free(m->v); free(m->t); free(m->fn); free(m->vn); free(m->en); free(m->te);
    free(m);
}

size_t ref_mesh_edges(const ref_mesh *m) { return m->ne; }

/* the closest point of triangle (a, b, c) to p and its feature: 0, 1, 2 the
 * vertices, 3 edge ab, 4 edge bc, 5 edge ca, 6 the face */
static int mesh_closest(const double *p, const double *a, const double *b, const double *c,
                        double q[3]) {
    double ab[3], ac[3], ap[3], bp[3], cp[3];
    sub3(b, a, ab);
    sub3(c, a, ac);
    sub3(p, a, ap);
    const double d1 = dot3(ab, ap), d2 = dot3(ac, ap);
    if (d1 <= 0.0 && d2 <= 0.0) {
        memcpy(q, a, 3 * sizeof(double));
        return 0;
    }
    sub3(p, b, bp);
    const double d3 = dot3(ab, bp), d4 = dot3(ac, bp);
    if (d3 >= 0.0 && d4 <= d3) {
        memcpy(q, b, 3 * sizeof(double));
        return 1;
    }
    const double vc = d1 * d4 - d3 * d2;
    if (vc <= 0.0 && d1 >= 0.0 && d3 <= 0.0) {
        const double s = d1 / (d1 - d3);
        for (int k = 0; k < 3; k++) q[k] = a[k] + s * ab[k];
        return 3;
    }
    sub3(p, c, cp);
    const double d5 = dot3(ab, cp), d6 = dot3(ac, cp);
    if (d6 >= 0.0 && d5 <= d6) {
        memcpy(q, c, 3 * sizeof(double));
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

double ref_mesh_signed_distance(const ref_mesh *m, const double p[3]) {
    double best = INFINITY, bq[3] = {0., 0., 0.};
    size_t bt = 0;
    int bf = -1;
    for (size_t f = 0; f < m->nt; f++) {
        double q[3], d[3];
        const int ft = mesh_closest(p, m->v + 3 * m->t[3 * f], m->v + 3 * m->t[3 * f + 1],
                                    m->v + 3 * m->t[3 * f + 2], q);
        sub3(p, q, d);
        const double d2 = dot3(d, d);
        if (d2 < best) {
            best = d2;
            bt = f;
            bf = ft;
            memcpy(bq, q, sizeof bq);
        }
    }
    if (bf < 0) return INFINITY;
    double diff[3];
    sub3(p, bq, diff);
    const double dist = norm3(diff);
    const double *n;
    if (bf < 3) n = m->vn + 3 * m->t[3 * bt + bf];
    else if (bf < 6) n = m->en + 3 * m->te[3 * bt + (bf - 3)];
    else n = m->fn + 3 * bt;
    return dot3(diff, n) < 0.0 ? -dist : dist;
}

void ref_mesh_sdf(const ref_mesh *m, const double *p, size_t np, double *out) {
#pragma omp parallel for schedule(dynamic, 64)
    for (size_t i = 0; i < np; i++) out[i] = ref_mesh_signed_distance(m, p + 3 * i);
}

void ref_mesh_sdf_build(const ref_mesh *m, const double x_min[3], const int64_t n[3], double dx,
                        double prune, double *values, uint8_t *present) {
    const int64_t nx = 3 * n[0] + 1, ny = 3 * n[1] + 1, nz = 3 * n[2] + 1;
    const size_t nodes = (size_t)(nx * ny * nz);
    uint8_t *isnode = (uint8_t *)calloc(nodes, 1);
    for (int64_t cz = 0; cz < n[2]; cz++)
        for (int64_t cy = 0; cy < n[1]; cy++)
            for (int64_t cx = 0; cx < n[0]; cx++)
                for (int i = 0; i < 32; i++)
                    isnode[((3 * cz + NODE_OFFS[i][2]) * ny + (3 * cy + NODE_OFFS[i][1])) * nx +
                           (3 * cx + NODE_OFFS[i][0])] = 1;
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t k = 0; k < (int64_t)nodes; k++) {
        values[k] = 0.0;
        present[k] = 0;
        if (!isnode[k]) continue;
        const int64_t ix = k % nx, iy = (k / nx) % ny, iz = k / (nx * ny);
        const double p[3] = {x_min[0] + (double)ix / 3.0 * dx, x_min[1] + (double)iy / 3.0 * dx,
                             x_min[2] + (double)iz / 3.0 * dx};
        values[k] = ref_mesh_signed_distance(m, p);
    }
    /* CubicSerendipityDiscretization::new pruning (discretization.rs:208-236):
     * a cell goes when all 32 node values lie below -prune or all above
     * +prune; nodes of kept cells are present */
    for (int64_t cz = 0; cz < n[2]; cz++)
        for (int64_t cy = 0; cy < n[1]; cy++)
            for (int64_t cx = 0; cx < n[0]; cx++) {
                int low = 1, high = 1;
                for (int i = 0; i < 32; i++) {
                    const double v = values[((3 * cz + NODE_OFFS[i][2]) * ny +
                                             (3 * cy + NODE_OFFS[i][1])) * nx +
                                            (3 * cx + NODE_OFFS[i][0])];
                    low &= v < -prune;
                    high &= v > prune;
                }
                if (low || high) continue;
                for (int i = 0; i < 32; i++)
                    present[((3 * cz + NODE_OFFS[i][2]) * ny + (3 * cy + NODE_OFFS[i][1])) * nx +
                            (3 * cx + NODE_OFFS[i][0])] = 1;
            }
    for (size_t k = 0; k < nodes; k++)
        if (!present[k]) values[k] = 0.0;
    free(isnode);
}
