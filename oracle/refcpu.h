/*
 * refcpu — CPU oracle for the SPH particle hot path.
 *
 * TEST INFRASTRUCTURE ONLY. This is a plain-C, f64 restatement of the
 * reference solver's per-step algorithm (SammyMcFly/fluid-simulation,
 * simulation_lib). It exists to check the HIP implementation and to serve as
 * the timed CPU baseline. Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it; the product path never does.
 *
 * Every function names the reference file:line it restates. Evaluation order
 * of floating-point expressions follows the reference's Rust expressions
 * (left-associative, powi(3) == (x*x)*x, nalgebra dot == (ax*bx+ay*by)+az*bz),
 * and the library is compiled with -ffp-contract=off, because rustc never
 * contracts a*b+c into an FMA.
 *
 * Parity pins (what checks each part) are listed in oracle/README.md.
 */
#ifndef REFCPU_H
#define REFCPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------- kernel (sph/kernel/mod.rs:17-46, cubic_spline.rs:8-17) */
double ref_w(double q);
double ref_dqw(double q);
double ref_kernel_function(const double r[3], double h);
void ref_kernel_gradient(const double r[3], double h, double out[3]);

/* ---------------- spatial hashing (neighbor_search/spatial_hashing.rs) */
void ref_get_cell(const double p[3], double cell_size, int32_t out[3]);
uint64_t ref_map_int_to_uint(int32_t k);
uint64_t ref_hash(int32_t k, int32_t l, int32_t m);

/* CSR neighbor list (neighbor_search/mod.rs:49-161). */
typedef struct ref_nlist {
    size_t n;          /* number of query samples */
    size_t total;      /* number of entries */
    uint64_t *offsets; /* n + 1 */
    uint64_t *indices; /* total */
} ref_nlist;

/* SpatialHashing::new(cell_size) + find_samples(range, positions, samples)
 * (spatial_hashing.rs:27-36, 75-103, 118-158). Allocates *out. */
int ref_find_samples(double cell_size, double range, const double *pos, size_t n,
                     const double *samples, size_t m, ref_nlist *out);
void ref_nlist_free(ref_nlist *nl);
/* helpers for ctypes callers: one call allocates and returns a handle */
ref_nlist *ref_find_samples_alloc(double cell_size, double range, const double *pos,
                                  size_t n, const double *samples, size_t m);
size_t ref_nlist_total(const ref_nlist *nl);
void ref_nlist_copy(const ref_nlist *nl, uint64_t *offsets, uint64_t *indices);
void ref_nlist_destroy(ref_nlist *nl);

/* ---------------- cubic serendipity lattice (utilities/discretization.rs) */
typedef struct ref_lattice {
    double x_min[3];
    double x_max[3];
    double dx;
    int64_t n[3];            /* cells per axis */
    const double *values;    /* dense nodes (3n+1)^3, x fastest */
    const uint8_t *present;  /* node present in the reference's HashMap */
} ref_lattice;

/* status codes of function()/gradient(): 0 ok, 1 OutOfBounds, 2 PrunedCell */
int ref_serendipity_function(const ref_lattice *l, const double p[3], double *out);
int ref_serendipity_gradient(const ref_lattice *l, const double p[3], double out[3]);
void ref_shape_functions(double xi, double eta, double zeta, double out[32]);
void ref_reference_nodes(double out[32 * 3]);
int ref_node_offset(double r); /* to_offset (discretization.rs:262-264) */

/* ---------------- rigid body (boundary_handling/rigid_body_motion.rs) */
typedef struct ref_rigid {
    double mass;
    double inertia_inv_body[9];  /* row-major */
    double com[3];
    double q[4];                 /* w, i, j, k */
    double lin_vel[3];
    double ang_mom[3];
    double force[3];
    double torque[3];
    double inertia_inv_world[9];
    double ang_vel[3];
} ref_rigid;

void ref_rigid_init(ref_rigid *rb, double mass, const double inertia_body[9],
                    const double inertia_inv_body[9], const double com[3],
                    const double q[4], const double lin_vel[3], const double ang_vel[3]);
void ref_rigid_add_force(ref_rigid *rb, const double f[3], const double loc[3]);
void ref_rigid_step(ref_rigid *rb, double dt);
void ref_rigid_world_to_local(const ref_rigid *rb, const double p[3], double out[3]);
void ref_rigid_local_to_world(const ref_rigid *rb, const double p[3], double out[3]);
void ref_rigid_rotate(const ref_rigid *rb, const double v[3], double out[3]);
void ref_rigid_velocity_at_point(const ref_rigid *rb, const double p[3], double out[3]);
void ref_quat_rotate(const double q[4], const double v[3], double out[3]);
void ref_quat_to_matrix(const double q[4], double R[9]);

/* ---------------- system (sph/mod.rs) */
enum { REF_SESPH = 0, REF_IISPH = 1, REF_IISPHWOST = 2, REF_SESPH_SPLITTING = 3 };
enum { REF_EULER_CROMER = 0, REF_TAKE_PREDICTED = 1, REF_EXPLICIT_EULER = 2, REF_VERLET = 3 };
enum { REF_GRAVITY_UNIFORM = 0, REF_GRAVITY_RADIAL = 1 };

typedef struct ref_params {
    double time_increment;      /* fixed dt (cfl == 0) */
    int32_t cfl;                /* 1: CFL adaptive dt (feature cfl_time_step) */
    double max_time_increment;
    double cfl_number;
    double kernel_support_radius;
    double rest_density_grid_spacing;
    double disable_particles_below;
    double fluid_viscosity;
    double boundary_viscosity;
    double boundary_pressure_acceleration_weighting;
    double boundary_rest_volume_weighting;
    int32_t gravity_mode;
    double gravity_center[3];
    int32_t solver;
    int32_t integrator;
    double stiffness;
    double target_density_error; /* percent */
    double relaxation_factor;
    double min_diagonal_element;
} ref_params;

typedef struct ref_measurement {
    double time;
    double time_step_size;
    double density;
    double density_error;
    double kinetic_energy;
    uint32_t solver_iterations;
    uint32_t eqs1_iterations;   /* IISPHwOST first solve, 0 otherwise */
    double predicted_density_error;
    uint64_t steps;
    uint64_t num_particles;
} ref_measurement;

typedef struct ref_system ref_system;

ref_system *ref_system_create(const ref_params *p, size_t n, const double *x,
                              const double *v, const double *m, const uint32_t *fluid_id);
void ref_system_destroy(ref_system *s);
/* StaticSampleBoundary entry: samples in world frame (static) or body frame
 * (dynamic; rigid_init != NULL). Returns boundary index or -1. */
int ref_system_add_static_samples(ref_system *s, const double *pos, size_t m,
                                  const double *rigid_init);
/* VolumeMapBoundary entry. rigid_init (or NULL) = mass, I_body[9], I_inv_body[9],
 * com[3], q[4], v[3], omega[3]  (32 doubles). */
int ref_system_add_volume_map(ref_system *s, const ref_lattice *sdf,
                              const ref_lattice *vm, const double *rigid_init);
/* The body-frame bounding box of boundary b's mesh (LocalAABB,
 * volume_map_boundary.rs:1137-1150, 1209-1222); a boundary without one takes
 * no part in rigid-rigid contact. */
int ref_system_set_aabb(ref_system *s, int b, const double center[3], const double half[3]);
/* Rigid-rigid contact of a system whose boundaries are all volume maps
 * (VolumeMapBoundary::step_forward_in_time, volume_map_boundary.rs:1440-1453):
 * the Gauss-Legendre rule (nodes ascending on [-1, 1]) of the contact-cell
 * quadrature, degree 8 in the reference (:57); degree 0 turns it off. */
int ref_system_set_contact(ref_system *s, const double *gl_x, const double *gl_w, int degree);
/* active contact cells and NNCG iterations of the last step */
void ref_system_contact_stats(const ref_system *s, uint32_t out[2]);
/* the last step's active contact cells in solve order, REF_CONTACT_REC
 * doubles each: body a, body b, volume, center[3], f_a[3], tau_a[3], f_b[3],
 * tau_b[3], a_cc pressure, a_cc friction, p, lambda[3]; returns the count */
#define REF_CONTACT_REC 24
size_t ref_system_contact_cells(const ref_system *s, double *out, size_t cap);
/* solve_for_p_c (volume_map_boundary.rs:186-637) after compute_diagonal, on
 * cell records as above (p and lambda are written) and REF_CONTACT_BODY
 * doubles per body: dynamic flag, mass, com[3], linear velocity[3], angular
 * momentum[3], force[3], torque[3], world inverse inertia[9] (row major).
 * extra[6 * nb] receives each body's committed contact force and torque;
 * returns the NNCG iteration count. */
#define REF_CONTACT_BODY 26
uint32_t ref_contact_solve(const double *bodies, int nb, double *cells, size_t nc, double dt,
                           double *extra);
/* BoundaryHandling::initialize + resize_slots + update() (System::new_boxed). */
int ref_system_init(ref_system *s);
int ref_system_step(ref_system *s, int nsteps);
size_t ref_system_len(const ref_system *s);
void ref_system_get(const ref_system *s, double *x, double *v, double *a, double *V,
                    double *p, double *m, uint32_t *fluid_id, uint64_t *origin);
void ref_system_measure(const ref_system *s, ref_measurement *out);
/* per-iteration predicted density errors of the last solve (TDE solve) */
size_t ref_system_error_history(const ref_system *s, double *out, size_t cap);
/* body state: com[3], q[4], lin_vel[3], ang_mom[3], force[3], torque[3], ang_vel[3] */
int ref_system_body_state(const ref_system *s, int b, double out[22]);
/* forces/torques applied to dynamic bodies during the last update(), in FP64:
 * F[3], tau[3] per boundary (zeros for static) */
int ref_system_body_forces(const ref_system *s, int b, double out[6]);
/* fluid neighbor list of the last update() */
const ref_nlist *ref_system_fluid_nlist(const ref_system *s);
/* boundary samples of boundary b for fluid particle i, mode 0 normal, 1 viscosity:
 * writes up to cap entries of (pos[3], vel[3], vol) and returns the count */
size_t ref_system_boundary_samples(const ref_system *s, int b, size_t i, int mode,
                                   double *out, size_t cap);
double ref_system_static_volume(const ref_system *s, int b, size_t k);
void ref_system_set_threads(int nthreads);

/* ---- volume-map construction (setup; DiscretizedBoundaryFields::new) ------
 * gauss_legendre_integrate_ball (utilities/discretization.rs:33-79) of
 * cubic_extension_fn(sdf(p), h) (volume_map_boundary.rs:1716-1726) around
 * `center`, with the Gauss-Legendre rule (gl_x, gl_w) of `degree` points on
 * [-1, 1] mapped to [a, b] as x -> ((b - a) x + (b + a)) / 2, sum times
 * (b - a) / 2, terms accumulated in node order. Returns 0, or the first
 * error status of an sdf evaluation (1 OutOfBounds, 2 PrunedCell). */
int ref_vm_integrate_ball(const ref_lattice *sdf, const double center[3], double h,
                          const double *gl_x, const double *gl_w, int degree, double *out);
/* SDFnWrapper::volume (volume_map_boundary.rs:1684-1714) clamped to
 * [0, ball_volume(h)] as in DiscretizedBoundaryFields::new (:1606-1624). */
int ref_vm_node_value(const ref_lattice *sdf, const double p[3], double h, const double *gl_x,
                      const double *gl_w, int degree, double *out);
/* The volume-map lattice over the cells of (x_min, n, dx)
 * (CubicSerendipityDiscretization::new, discretization.rs:99-245, prune lower
 * bound 0): values[] and present[] of the dense (3n+1)^3 nodes. A cell is kept
 * iff all 32 of its node values evaluate; present = node of a kept cell.
 * Returns the number of nodes that ran the ball quadrature. */
uint64_t ref_vm_build(const ref_lattice *sdf, double h, const double x_min[3], const int64_t n[3],
                  double dx, const double *gl_x, const double *gl_w, int degree, double *values,
                  uint8_t *present);

/* SPHSystem::get_quantity_at_positions (sph/mod.rs:241-341): quantity
 * 0 Speed, 1 Volume, 2 Density, 3 DensityError, 4 Pressure, 5 KineticEnergy
 * (render_info.rs:205-212) at nq f32 positions; out = f32 values. */
int ref_system_quantity_at_positions(ref_system *s, int quantity, size_t nq, const float *pos,
                                     float *out);

/* ---- signed distance to an oriented triangle mesh ---------------------------
 * TriangleMeshWrapper::signed_distance (volume_map_boundary.rs:1655-1671):
 * the distance to the mesh's closest point, negative where the offset points
 * against the outward pseudo-normal of the closest feature (face, edge or
 * vertex). vertices [nv][3], triangles [nt][3] with outward (counter-
 * clockwise) winding and shared vertices already merged; NULL on an index
 * out of range. */
typedef struct ref_mesh ref_mesh;
ref_mesh *ref_mesh_create(const double *vertices, size_t nv, const uint32_t *triangles,
                          size_t nt);
void ref_mesh_destroy(ref_mesh *m);
size_t ref_mesh_edges(const ref_mesh *m);
double ref_mesh_signed_distance(const ref_mesh *m, const double p[3]);
void ref_mesh_sdf(const ref_mesh *m, const double *p, size_t np, double *out);
/* The signed-distance lattice of DiscretizedBoundaryFields::new (:1547-1591)
 * over (x_min, n, dx), pruned as CubicSerendipityDiscretization::new does
 * with bounds -prune / +prune (discretization.rs:208-236). */
void ref_mesh_sdf_build(const ref_mesh *m, const double x_min[3], const int64_t n[3], double dx,
                        double prune, double *values, uint8_t *present);

/* ---- test hooks ---------------------------------------------------------
 * One reference function on the system's current state (after
 * ref_system_init), so per-function tests can set volumes, pressures and
 * predicted velocities directly:
 *   NEIGHBORS            fluid neighbor list + boundary samples (update())
 *   VOLUME               get_volume (quantities.rs:14-61)
 *   RESET_ACCELERATION   reset_acceleration (non_pressure_accelerations.rs:19-27)
 *   GRAVITY, VISCOSITY   (non_pressure_accelerations.rs:64-185)
 *   PRESSURE_ACCELERATION[_OVERWRITE]  add_pressure_acceleration into a, with
 *                        reaction forces (pressure_solver/mod.rs:121-195);
 *   PRESSURE_ACCELERATION_SCRATCH      the same into the IISPH scratch, no reaction
 *   SESPH                SESPH::solve_and_add_acceleration (sesph.rs:31-67)
 *   PRED_VEL             set_pred_vel_by_applying_acc(v) (pressure_solver/mod.rs:90-108)
 *   SOURCE_VDE/VP        IISPH source terms (iisph.rs:117-251)
 *   DIAGONAL, INITIALIZE (iisph.rs:254-376)
 *   RESOLVE              resolve_pressure_globally with the target density
 *                        error; returns the iteration count */
enum {
    REF_FN_NEIGHBORS = 0, REF_FN_VOLUME, REF_FN_RESET_ACCELERATION, REF_FN_GRAVITY,
    REF_FN_VISCOSITY, REF_FN_PRESSURE_ACCELERATION, REF_FN_PRESSURE_ACCELERATION_OVERWRITE,
    REF_FN_PRESSURE_ACCELERATION_SCRATCH, REF_FN_SESPH, REF_FN_PRED_VEL, REF_FN_SOURCE_VDE,
    REF_FN_SOURCE_VP, REF_FN_DIAGONAL, REF_FN_INITIALIZE, REF_FN_RESOLVE,
    REF_FN_SESPH_SPLITTING, REF_FN_INTEGRATE
};
int ref_system_call(ref_system *s, int fn);
/* read (write = 0) or overwrite (write = 1) one field; returns its length in
 * doubles: x, v, a [n][3]; m, V, p [n]; predicted velocity / position [n][3];
 * source term s_f, diagonal a_ii [n]; pressure-acceleration scratch [n][3];
 * time step (1) */
enum {
    REF_F_X = 0, REF_F_VEL, REF_F_ACC, REF_F_MASS, REF_F_VOLUME, REF_F_PRESSURE, REF_F_PRED_VEL,
    REF_F_PRED_POS, REF_F_SOURCE, REF_F_DIAGONAL, REF_F_PRESSURE_ACC, REF_F_TIME_STEP,
    REF_F_SLOT_POS /* the integrator's position slot (TakePredicted, Verlet) */
};
int ref_system_field(ref_system *s, int field, int write, double *buf);
/* pseudo volumes of static-sample boundary b (read or overwrite) */
int ref_system_sample_volumes(ref_system *s, int b, int write, double *buf);

#ifdef __cplusplus
}
#endif
#endif
