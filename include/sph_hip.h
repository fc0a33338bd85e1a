/*
 * sph_hip.h — C ABI of the MI355X (gfx950) SPH particle hot path.
 *
 * The entry points replace the three plug-in traits of the reference solver
 * (SammyMcFly/fluid-simulation, simulation_lib) that System<K,I,P,N,B> is
 * generic over, plus a device-resident stepping path:
 *
 *   NeighborSearch     neighbor_search/mod.rs:18-28
 *   PressureSolver     sph/pressure_solver/mod.rs:37-78
 *   BoundaryHandling   sph/boundary_handling/mod.rs:39-119
 *   System::step_forward_in_time / update   sph/mod.rs:125-144, 674-731
 *
 * Conventions
 *   - every function returns 0 on success or a negative sph_status; no C++
 *     exception crosses the ABI; sph_last_error() returns a thread-local
 *     message for the last failure on the calling thread;
 *   - host arrays are caller-owned and only read/written during the call;
 *     device buffers are owned by the context;
 *   - positions/velocities are [n][3] doubles (the layout of
 *     &[Point3<f64>] / &[Vector3<f64>] in the reference); neighbor indices are
 *     uint64_t (usize in the reference);
 *   - a context is used by one thread at a time (the reference's Send, not
 *     Sync); independent contexts run on their own HIP streams;
 *   - all arithmetic is IEEE binary64 with no FMA contraction.
 */
#ifndef SPH_HIP_H
#define SPH_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum sph_status {
    SPH_OK = 0,
    SPH_ERR_ARG = -1,          /* invalid argument */
    SPH_ERR_HIP = -2,          /* HIP runtime error (device, launch, memory) */
    SPH_ERR_STATE = -3,        /* call out of order (e.g. step before init) */
    SPH_ERR_CAPACITY = -4,     /* caller buffer too small; *total tells the size */
    SPH_ERR_UNSUPPORTED = -5,  /* input outside what this build handles */
    SPH_ERR_NO_DEVICE = -6     /* no GPU visible */
} sph_status;

enum { SPH_SOLVER_SESPH = 0, SPH_SOLVER_IISPH = 1, SPH_SOLVER_IISPHWOST = 2,
       SPH_SOLVER_SESPH_SPLITTING = 3 };
enum { SPH_INTEGRATOR_EULER_CROMER = 0, SPH_INTEGRATOR_TAKE_PREDICTED = 1,
       SPH_INTEGRATOR_EXPLICIT_EULER = 2, SPH_INTEGRATOR_VERLET = 3 };
enum { SPH_GRAVITY_UNIFORM = 0, SPH_GRAVITY_RADIAL = 1 };

/* [parameters] + [procedures] of the reference (sph/setup/input.rs:103-260,
 * SystemParameters sph/mod.rs:755-830). */
typedef struct sph_params {
    double time_increment;           /* fixed dt when cfl == 0 */
    int32_t cfl;                     /* 1: CFL adaptive dt (feature cfl_time_step) */
    double max_time_increment;
    double cfl_number;
    double kernel_support_radius;    /* h */
    double rest_density_grid_spacing;/* dx */
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
    double target_density_error;     /* percent */
    double relaxation_factor;
    double min_diagonal_element;
} sph_params;

/* CubicSerendipityDiscretization as a dense node lattice
 * (utilities/discretization.rs:84-92): (3n+1)^3 nodes, x fastest;
 * present[k] != 0 where the reference's HashMap holds node k. */
typedef struct sph_lattice {
    double x_min[3];
    double x_max[3];
    double dx;
    int64_t n[3];
    const double *values;
    const uint8_t *present;
} sph_lattice;

/* RigidBodyMotion::new arguments (boundary_handling/rigid_body_motion.rs:38-62);
 * matrices row-major, quaternion (w, i, j, k). */
typedef struct sph_rigid_init {
    double mass;
    double inertia_body[9];
    double inertia_inv_body[9];
    double com[3];
    double q[4];
    double lin_vel[3];
    double ang_vel[3];
} sph_rigid_init;

/* Measurement row (measurement.rs:52-75) plus step counters. */
typedef struct sph_measurement {
    double time;
    double time_step_size;
    double density;                  /* CurrentSystemProperties::average_density */
    double density_error;            /* percent, sph/mod.rs:489-529 */
    double kinetic_energy;
    uint32_t solver_iterations;      /* last solve (EQS2 for IISPHwOST) */
    uint32_t eqs1_iterations;        /* IISPHwOST EQS1, else 0 */
    double predicted_density_error;  /* percent */
    uint64_t steps;
    uint64_t num_particles;
} sph_measurement;

const char *sph_last_error(void);

/* ---- lifecycle ------------------------------------------------------- */
typedef struct sph_ctx sph_ctx;
int sph_ctx_create(const sph_params *params, sph_ctx **out);
/* Deep copy of all device state into an independent context
 * (SPHSystem: DynClone, sph/mod.rs:40,71). */
int sph_ctx_clone(const sph_ctx *src, sph_ctx **out);
int sph_ctx_destroy(sph_ctx *ctx);

/* ---- NeighborSearch::find_samples (neighbor_search/mod.rs:18-28,
 *      spatial_hashing.rs:130-158) ------------------------------------------
 * SpatialHashing with cell size `cell_size`: for each of the n query
 * positions, the indices of the m samples closer than `range` (strict),
 * in the reference's order. offsets has n+1 entries; indices has room for
 * `cap` entries; *total receives the entry count (SPH_ERR_CAPACITY if it
 * exceeds cap). Where two cells of a query's stencil share a spatial-hash
 * bucket (equal 64-bit keys), the bucket's samples in range are listed at
 * each of the two cells, as the reference's HashMap lists them. */
int sph_neighbors_find(sph_ctx *ctx, double cell_size, double range, const double *pos,
                       size_t n, const double *samples, size_t m, uint64_t *offsets,
                       uint64_t *indices, size_t cap, size_t *total);

/* Allocation-free variant of sph_neighbors_find. No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * pos, samples, offsets and indices are DEVICE pointers ([n][3] and [m][3]
 * doubles, n+1 and cap uint64_t); *total is host memory. The results equal
 * sph_neighbors_find's. It calls no hipMalloc/hipFree: all intermediate
 * buffers live in a static device array of 25 MiB per GPU, held from the
 * library's load until the process ends. Limits: n, m <= 262144 and an
 * occupied cell box of the samples of at most 2^20 cells, else
 * SPH_ERR_UNSUPPORTED before any launch that writes the outputs; inputs whose
 * stencils hold cells that share a spatial-hash bucket are refused with
 * SPH_ERR_UNSUPPORTED too (the array holds no table of shared buckets, which
 * sph_neighbors_find builds). Every call on
 * one GPU shares that array: calls that overlap on one GPU (several threads,
 * several contexts) race. The call returns after its work on the context's
 * stream has finished; the caller makes the inputs visible to that stream
 * first (e.g. a device synchronization). */
int sph_neighbors_find_noalloc(sph_ctx *ctx, double cell_size, double range, const double *pos,
                               size_t n, const double *samples, size_t m, uint64_t *offsets,
                               uint64_t *indices, size_t cap, size_t *total);

/* ---- BoundaryHandling::add_static_boundary / add_dynamic_boundary ----------
 * (boundary_handling/mod.rs:44-60). StaticSampleBoundary: sample positions
 * (world frame when rigid == NULL, body frame otherwise: the samples then
 * follow the body's rigid motion, static_sample_boundary.rs:79-130; their
 * pseudo volumes are those of the initial pose).
 * VolumeMapBoundary: signed-distance and volume-map lattices (body frame for
 * dynamic bodies). Return the boundary index (>= 0) or a negative status. */
int sph_boundary_add_samples(sph_ctx *ctx, const double *pos, size_t m,
                             const sph_rigid_init *rigid);
int sph_boundary_add_volume_map(sph_ctx *ctx, const sph_lattice *sdf, const sph_lattice *vm,
                                const sph_rigid_init *rigid);

/* ---- rigid-rigid contact (VolumeMapBoundary::step_forward_in_time,
 * volume_map_boundary.rs:1440-1453) -----------------------------------------
 * Before the bodies of a volume-map system step, the overlapping world AABBs
 * of every body pair with a dynamic member are cut into sub-cells of h/2
 * (generate_contact_cells, :98-161); Gauss-Legendre quadrature of the
 * cubic-extended signed distances gives each cell's contact volume and
 * sensitivities (:702-946), and an NNCG-accelerated projected Jacobi with
 * Coulomb friction (solve_for_p_c, :186-637) adds the contact forces and
 * torques to the dynamic bodies.
 * The body-frame bounding box of boundary b's mesh (LocalAABB, :1137-1150 and
 * :1209-1222: parry's local_aabb center and half extents); a boundary without
 * one takes no part in contact. */
int sph_boundary_set_aabb(sph_ctx *ctx, int b, const double center[3],
                          const double half_extents[3]);
/* The contact-cell quadrature rule: degree Gauss-Legendre nodes (ascending on
 * [-1, 1]) and weights, degree 8 in the reference (:57); 1..16, or 0 to turn
 * contact off (the default). */
int sph_contact_set_quadrature(sph_ctx *ctx, const double *gl_x, const double *gl_w,
                               int degree);
/* active contact cells and NNCG iterations of the last step */
int sph_contact_stats(sph_ctx *ctx, uint32_t out[2]);
/* The last step's active cells in solve order, 24 binary64 each: body a,
 * body b, volume, center[3], f_a[3], tau_a[3], f_b[3], tau_b[3], a_cc
 * pressure, a_cc friction, p, lambda[3]; *count receives the cell count
 * (SPH_ERR_CAPACITY if it exceeds cap). */
int sph_contact_cells(sph_ctx *ctx, double *out, size_t cap, size_t *count);

/* ---- device-resident system ----------------------------------------------- */
/* Fluid::add_samples equivalent: the fluid state (sph/fluid.rs:26-104). */
int sph_fluid_upload(sph_ctx *ctx, size_t n, const double *x, const double *v, const double *m,
                     const uint32_t *fluid_id);
/* BoundaryHandling::initialize + System::new_boxed's update() (sph/mod.rs:389-421). */
int sph_system_init(sph_ctx *ctx);
/* nsteps x SPHSystem::step_forward_in_time (sph/mod.rs:125-144); *out may be NULL. */
int sph_step(sph_ctx *ctx, int nsteps, sph_measurement *out);
int sph_fluid_len(const sph_ctx *ctx, size_t *n);
/* Fluid state in the reference's particle order (any pointer may be NULL). */
int sph_fluid_download(sph_ctx *ctx, double *x, double *v, double *a, double *V, double *p,
                       double *m, uint32_t *fluid_id);
/* SPHSystem::take_measurement (sph/mod.rs:147-184). */
int sph_measure(sph_ctx *ctx, sph_measurement *out);
/* predicted density error (%) of every Jacobi iteration of the last solve */
int sph_error_history(sph_ctx *ctx, double *out, size_t cap, size_t *count);
/* rigid state: com[3], q[4], lin_vel[3], ang_mom[3], force[3], torque[3], ang_vel[3] */
int sph_body_state(sph_ctx *ctx, int b, double out[22]);
/* F[3], tau[3] applied to boundary b during the last update() (FP64) */
int sph_body_forces(sph_ctx *ctx, int b, double out[6]);

/* ---- checkpoints ------------------------------------------------------------------
 * The integrator's position and velocity slots of FluidCheckpoint
 * (fluid.rs:352-381, IntegrationScheme::POSITION_SLOTS / VELOCITY_SLOTS,
 * integration_schemes/mod.rs:23-42): one position and one velocity slot for
 * TakePredicted (take_predicted.rs:9-10), one position slot for Verlet (the
 * position one step back, verlet.rs:10-11), none for EulerCromer and
 * ExplicitEuler. sph_fluid_slot_counts reports them; sph_fluid_download_slots
 * sets *nslots to the position-slot count and fills xp and vp (either may be
 * NULL; vp only where there is a velocity slot) with [n][3] in the
 * reference's particle order. */
int sph_fluid_slot_counts(sph_ctx *ctx, int *position_slots, int *velocity_slots);
int sph_fluid_download_slots(sph_ctx *ctx, int *nslots, double *xp, double *vp);
/* Sets the integrator's slots ([n][3] in the reference's particle order;
 * either may be NULL; ignored where the scheme has no such slot). No
 * counterpart in the reference, whose slots always start where
 * Fluid::resize_slots leaves them (fluid.rs:112-129): Verlet's position slot
 * at the origin, so its first step moves each particle to about twice its
 * position. This lets a caller start Verlet from rest (the slot at the
 * positions) instead. sph_system_restore resets the slots, as the
 * reference's resume does. */
int sph_fluid_upload_slots(sph_ctx *ctx, const double *xp, const double *vp);
/* SPHSystem::continue_from_checkpoint (sph/mod.rs:362-378), called after
 * sph_fluid_upload of the checkpoint's fluid and in place of sph_system_init:
 * restores the step count and current time, each dynamic boundary's
 * RigidBodyMotion state (rigid_body_motion.rs:214-222), re-sizes the
 * integrator's slots as Fluid::resize_slots does (origin and zero: the
 * reference does not read a checkpoint's saved slots back) and runs update().
 * bodies holds 19 binary64 per boundary: center of mass[3], orientation
 * (w, i, j, k; normalized as UnitQuaternion::from_quaternion does,
 * rigid_body_motion.rs:273-289), linear velocity[3], angular momentum[3],
 * force[3], torque[3]; present[b] = 0 is a checkpoint entry of None. As in
 * BoundaryHandling::restore_from_checkpoint (volume_map_boundary.rs:
 * 1504-1518, 2099-2120), a count mismatch skips every boundary and a kind
 * mismatch skips that boundary. */
int sph_system_restore(sph_ctx *ctx, uint64_t time_steps_propagated, double current_time,
                       size_t nbodies, const uint8_t *present, const double *bodies);

/* ---- kernel timing ------------------------------------------------------------
 * HIP-event timing of the hot kernels on the context's stream (no counterpart
 * in the reference, which only times whole steps, sph/mod.rs:127,143).
 * sph_profile(ctx, 1) clears and starts; sph_profile_read returns the launch
 * count, the particles processed and the summed kernel milliseconds. */
enum {
    SPH_PROF_NEIGHBOR_COUNT = 0,   /* cell sweep: neighbor counts (one-sweep path: + pair cache + volume) */
    SPH_PROF_PAIRS_VOLUME = 1,     /* cell sweep: pair cache (index, grad W) + volume */
    SPH_PROF_BOUNDARY_SAMPLE = 2,  /* volume-map boundary samples */
    SPH_PROF_NONPRESSURE = 3,      /* gravity + viscosity */
    SPH_PROF_SOURCE_DIAG = 4,      /* IISPH source term + diagonal + initial pressure */
    SPH_PROF_PRESSURE_ACCEL = 5,   /* pressure acceleration */
    SPH_PROF_JACOBI = 6,           /* relaxed Jacobi update + error */
    SPH_PROF_VM_QUADRATURE = 7,    /* volume-map ball quadrature (units: band nodes) */
    SPH_PROF_QUANTITY = 8,         /* sensor queries (units: positions) */
    SPH_PROF_MESH_SDF = 9,         /* mesh signed distance at lattice nodes (units: nodes) */
    SPH_PROF_CONTACT = 10,         /* rigid-rigid contact cells (units: candidate sub-cells) */
    SPH_PROF_KERNELS = 11
};
int sph_profile(sph_ctx *ctx, int enable);
int sph_profile_read(sph_ctx *ctx, int kernel, uint64_t *launches, uint64_t *units, double *ms);

/* ---- per-trait entry points on host arrays --------------------------------
 * Each call uploads the given fluid state, runs one phase of System::update
 * on the device with the context's boundaries, and returns the result in the
 * caller's particle order. */
/* get_volume (sph/quantities.rs:14-61) */
int sph_volume(sph_ctx *ctx, size_t n, const double *x, double *V_out);
/* VolumeMapBoundary::find_boundary_samples for boundary b
 * (volume_map_boundary.rs:1259-1424): per particle, 7 doubles (pos[3], vel[3],
 * vol) of the normal-mode sample and flags[i] = 1 when it exists. */
int sph_boundary_find_samples(sph_ctx *ctx, int b, size_t n, const double *x,
                              double *samples_out, uint8_t *flags_out);
/* reset_acceleration + add_gravity_acceleration + add_viscosity_acceleration
 * (sph/non_pressure_accelerations.rs:19-185); V must hold get_volume's result. */
int sph_nonpressure_accel(sph_ctx *ctx, size_t n, const double *x, const double *v,
                          const double *m, const double *V, double *a_out,
                          double *body_ft_out);
/* PressureSolver::solve_and_add_acceleration (pressure_solver/mod.rs:65-72)
 * for params.solver: a_inout holds the non-pressure acceleration on entry and
 * the total on return; p_out the pressure; body_ft_out (6 per boundary, may be
 * NULL) the force/torque mirrored onto dynamic bodies; for IISPHwOST,
 * xpred_out/vpred_out (may be NULL) the predicted state committed by
 * TakePredicted. */
int sph_pressure_solve(sph_ctx *ctx, size_t n, const double *x, const double *v,
                       const double *m, const double *V, double dt, double *a_inout,
                       double *p_out, double *body_ft_out, double *xpred_out,
                       double *vpred_out, sph_measurement *info_out);

/* ---- worlds: many systems advanced in shared launches ------------------------
 * The reference steps one System at a time on its worker thread
 * (worker_loop, sci-phi-backend/src/lib.rs:507, a thread spawned by
 * sci-phi/src/app/mod.rs:236). A world
 * advances every attached context by whole time steps in one sequence of
 * multi-system launches: each launch covers all systems' particles, and every
 * per-system result (state, step size, iteration counts, error history,
 * measurements) is bit-identical to stepping that context alone with
 * sph_step. Each system keeps its own time step size and pressure-solve
 * termination; the host waits once per Jacobi iteration for all systems.
 * All systems of a world use one pressure solver (params.solver).
 * An attached context is stepped only through its world; sph_measure,
 * sph_fluid_download, sph_error_history and sph_body_* still read it, and
 * sph_step / the per-trait calls fail with SPH_ERR_STATE until it is detached.
 * A world and its contexts are used by one thread at a time. */
typedef struct sph_world sph_world;
int sph_world_create(sph_world **out);
/* detaches every context (they stay valid) and frees the world */
int sph_world_destroy(sph_world *world);
/* ctx must have run sph_system_init */
int sph_world_attach(sph_world *world, sph_ctx *ctx);
int sph_world_detach(sph_world *world, sph_ctx *ctx);
/* SPHSystem::step_forward_in_time for every attached system, nsteps times.
 * If a system's step fails (see sph_system_status), the call returns its
 * status after completing that step for every other system, which stays
 * valid; the failed system should be detached. */
int sph_world_step(sph_world *world, int nsteps);
/* 0, or the status with which the context's last step failed (the message
 * becomes sph_last_error()); its state is then not valid. */
int sph_system_status(const sph_ctx *ctx, int *status);
int sph_world_size(const sph_world *world, size_t *systems, size_t *particles);
/* kernel timing of the world's launches (as sph_profile / sph_profile_read) */
int sph_world_profile(sph_world *world, int enable);
int sph_world_profile_read(sph_world *world, int kernel, uint64_t *launches,
                           uint64_t *units, double *ms);
/* ---- serving: continuous batching over a world ----------------------------------
 * No counterpart in the reference, whose worker thread steps one system at a
 * time (worker_loop, sci-phi-backend/src/lib.rs:507). A server admits requests in submission
 * order while a slot and the HBM budget allow, keeps admitted systems resident
 * in one world per pressure solver and steps them together, and hands back
 * each system once it has taken its requested steps.
 * Protocol, per scheduling round: sph_server_submit for arrived requests
 * (a device-memory estimate and a step count; nothing is allocated);
 * sph_server_admit while it returns 1, building each admitted request's
 * context only then (so allocations follow admission) and handing it over
 * with sph_server_attach; sph_server_step (one step of every world, then finished
 * systems leave the world); sph_server_poll while it returns 1, which returns
 * a finished context to the caller and releases its slot and budget.
 * Creating a server and the submit/admit bookkeeping touch no device; the
 * worlds are created at the first attach of a system of their solver. One
 * thread at a time. */
typedef struct sph_server sph_server;
typedef struct sph_server_stats {
    uint64_t queued;          /* submitted, not yet admitted */
    uint64_t admitted;        /* admitted and not yet polled (slots in use) */
    uint64_t resident;        /* attached to the world */
    uint64_t finished;        /* finished, waiting for sph_server_poll */
    uint64_t bytes;           /* estimates of the admitted requests */
    uint64_t world_steps;     /* world steps taken */
    uint64_t failed;          /* failed, waiting for sph_server_poll_failed */
    uint64_t device_bytes;    /* device memory the library holds (sph_device_bytes_held) */
} sph_server_stats;
int sph_server_create(uint64_t budget_bytes, int max_slots, sph_server **out);
/* frees the worlds; contexts still attached are detached, not destroyed */
int sph_server_destroy(sph_server *server);
/* SPH_ERR_ARG if device_bytes exceeds the whole budget or steps is 0 */
int sph_server_submit(sph_server *server, uint64_t device_bytes, uint64_t steps,
                      uint64_t *ticket);
/* returns 1 and the oldest queued ticket if it fits (taking its slot and
 * bytes), 0 if the queue is empty or its head does not fit */
int sph_server_admit(sph_server *server, uint64_t *ticket);
/* ctx (after sph_system_init) serves an admitted ticket; it joins the world
 * of its pressure solver */
int sph_server_attach(sph_server *server, uint64_t ticket, sph_ctx *ctx);
/* gives up an admitted ticket whose system could not be built (e.g. its
 * System::new update failed), releasing its slot and bytes */
int sph_server_cancel(sph_server *server, uint64_t ticket);
/* one step of every world that holds a resident system; systems that have
 * taken their steps leave their world for the finished queue. A system whose
 * step fails (its pressure solve does not meet the termination test within
 * its SPH_OPT_MAX_ITERATIONS cap) leaves for the failed queue and the others
 * complete the step (the call returns SPH_OK); an error not confined to one
 * system (a HIP error) moves every system of that world to the failed queue,
 * the other worlds complete the step, and the error is returned. *out may be
 * NULL. */
int sph_server_step(sph_server *server, sph_server_stats *out);
/* returns 1 with the oldest failed ticket, its context (detached, now the
 * caller's, its state not valid) and the failure status, releasing its slot
 * and bytes; 0 if none has failed. No counterpart in the reference, whose
 * solve loops without a cap. */
int sph_server_poll_failed(sph_server *server, uint64_t *ticket, sph_ctx **ctx, int *status);
/* returns 1 with the oldest finished ticket and its context (detached, now
 * the caller's), releasing its slot and bytes; 0 if none has finished */
int sph_server_poll(sph_server *server, uint64_t *ticket, sph_ctx **ctx);
int sph_server_read_stats(const sph_server *server, sph_server_stats *out);
/* an implementation option (SPH_OPT_*) of the server's worlds, present and
 * future */
int sph_server_set_option(sph_server *server, int option, int value);
/* ---- sensor queries -------------------------------------------------------------
 * SPHSystem::get_quantity_at_positions (sph/mod.rs:241-341): a scalar field
 * of the current fluid state at nq positions ([nq][3] f32, widened to f64),
 * each an SPH sum over the fluid neighbors of the position within the kernel
 * support (and, for the speed and the volume, the boundaries' samples of the
 * position), in the reference's neighbor order; out: nq f32 values. The
 * quantities are those of quantities.rs: get_speed (:64-120), get_volume
 * (:14-61), get_density (:123-170), get_density_error (:173-228, evaluated
 * over the fluid masses, which the reference passes as its sample volumes),
 * get_pressure (:231-280), get_kinetic_energy (:283-330). The system's state
 * is not changed. */
enum {
    SPH_QUANTITY_SPEED = 0,  /* ScalarQuantity order (render_info.rs:205-212) */
    SPH_QUANTITY_VOLUME = 1,
    SPH_QUANTITY_DENSITY = 2,
    SPH_QUANTITY_DENSITY_ERROR = 3,
    SPH_QUANTITY_PRESSURE = 4,
    SPH_QUANTITY_KINETIC_ENERGY = 5
};
int sph_quantity_at_positions(sph_ctx *ctx, int quantity, size_t nq, const float *pos,
                              float *out);

/* ---- boundary-field construction (setup of VolumeMapBoundary) ----------------
 * The volume map of DiscretizedBoundaryFields::new (volume_map_boundary.rs:
 * 1593-1624) over the cells (x_min, n, dx): for every serendipity node p,
 * SDFnWrapper::volume (:1684-1714) clamped to [0, ball_volume(h)], i.e. 0
 * where `sdf` cannot be evaluated at p or sdf(p) >= 2h, the ball volume where
 * sdf(p) <= -h, else gauss_legendre_integrate_ball (utilities/discretization.rs:
 * 33-79) of cubic_extension_fn(sdf, h) (:1716-1726) with the `degree`-point
 * Gauss-Legendre rule (gl_x, gl_w) on [-1, 1] (each of the r, theta, phi rules
 * maps x to ((b - a) x + (b + a)) / 2 and scales its node-order sum by
 * (b - a) / 2). A node whose quadrature hits an unevaluable sdf point has no
 * value; a cell is kept iff all 32 of its nodes have values
 * (CubicSerendipityDiscretization::new, discretization.rs:99-245, lower bound
 * 0). Outputs: values and present flags of the dense (3n+1)^3 nodes, x
 * fastest. degree <= 32. `ctx` may be NULL (the build then uses a stream of
 * its own); otherwise it runs on the context's stream. `quadrature_nodes`
 * (may be NULL) receives the number of nodes that ran the ball quadrature. */
int sph_volume_map_build(sph_ctx *ctx, const sph_lattice *sdf, double h, const double x_min[3],
                         const int64_t n[3], double dx, const double *gl_x, const double *gl_w,
                         int degree, double *values_out, uint8_t *present_out,
                         uint64_t *quadrature_nodes);
/* The signed-distance half of DiscretizedBoundaryFields::new (volume_map_
 * boundary.rs:1547-1591) over the cells (x_min, n, dx): at every serendipity
 * node, TriangleMeshWrapper::signed_distance (:1655-1671) to the triangle
 * mesh (vertices [nv][3], triangles [nt][3] of vertex indices, outward
 * counter-clockwise winding, shared vertices merged): the distance to the
 * closest point of the first closest triangle, negative where the offset
 * points against the outward pseudo-normal of that point's feature (face;
 * edge: the normalized sum of its faces' normals; vertex: the angle-weighted
 * normalized sum). Cells whose 32 node values all lie below -prune or all
 * above +prune are pruned (CubicSerendipityDiscretization::new,
 * discretization.rs:208-236; the reference prunes at 3.1 h). Outputs as for
 * sph_volume_map_build; `ctx` may be NULL. */
int sph_mesh_sdf_build(sph_ctx *ctx, const double *vertices, size_t nv,
                       const uint32_t *triangles, size_t nt, const double x_min[3],
                       const int64_t n[3], double dx, double prune, double *values_out,
                       uint8_t *present_out);

/* free and total device memory (hipMemGetInfo), for admission control */
int sph_device_memory(size_t *free_bytes, size_t *total_bytes);
/* device memory the library's buffers hold right now, over every context,
 * world and server (no counterpart in the reference: for admission control,
 * beside the per-request estimates a server is given) */
int sph_device_bytes_held(uint64_t *bytes);

/* ---- implementation options (no counterpart in the reference) ----------------
 * None of them changes a result bit; they trade memory and speed.
 *   SPH_OPT_CACHE_GRADIENTS  1 = the neighbor passes read grad W from the pair
 *       cache (24 B more HBM per pair slot), 0 = they recompute it (default).
 *   SPH_OPT_AUTOTUNE  steps per code shape of the online tuner (default 1;
 *       0 = off): the first steps of a context or world, and the steps after
 *       its particle count changed 2x, time each code shape (unroll, waves per
 *       SIMD, loop form: plain, software-pipelined, two neighbors per step,
 *       grouped gathers) of the five neighbor passes, which then keep their
 *       fastest.
 *   SPH_OPT_SHAPE  (set) run every neighbor pass in this code shape
 *       (0 .. SPH_SHAPES-1) and turn the tuner off.
 *   SPH_OPT_TUNE_STATE  (get) 0 untuned, 1 measuring, 2 tuned.
 *   SPH_OPT_SHAPE_VOLUME .. SPH_OPT_SHAPE_VOLUME+4  (get) the code shape of the
 *       volume, non-pressure, source/diagonal, pressure-acceleration and
 *       Jacobi passes; (set) run that pass in this shape and turn the tuner
 *       off (the other passes keep theirs).
 *   SPH_OPT_GRAPHS  1 = replay repeated Jacobi iteration chunks from HIP
 *       graphs (default), 0 = launch every kernel directly. The replayed
 *       kernels read the batch layout from device memory, so a world's graphs
 *       serve every admission and departure.
 *   SPH_OPT_GRAPH_COUNT  (get) executable graphs the context or world holds.
 *   SPH_OPT_MAX_ITERATIONS  (context) Jacobi iterations after which a pressure
 *       solve that has not met its termination test fails with
 *       SPH_ERR_UNSUPPORTED (default 100000; the reference has no cap,
 *       iisph.rs:387-390).
 *   SPH_OPT_SWEEP  1 = neighbor list and volume from one cell sweep into a
 *       fixed-width pair cache (default), 0 = separate count, fill and
 *       volume passes into a slice-width pair cache. The pair cache with
 *       stored grad W (SPH_OPT_CACHE_GRADIENTS) always uses the passes. A
 *       value above 1 also sets the width of the next sweep's cache; a
 *       sweep that finds more neighbors than fit runs again, wider.
 * The options of a context apply when it steps alone, those of a world to
 * the world's steps. */
enum {
    SPH_OPT_CACHE_GRADIENTS = 0,
    SPH_OPT_AUTOTUNE = 1,
    SPH_OPT_SHAPE = 2,
    SPH_OPT_TUNE_STATE = 3,
    SPH_OPT_SHAPE_VOLUME = 4,
    SPH_OPT_GRAPHS = 9,
    SPH_OPT_SWEEP = 10,
    SPH_OPT_MAX_ITERATIONS = 11,
    SPH_OPT_GRAPH_COUNT = 12
};
enum { SPH_SHAPES = 9 };
int sph_set_option(sph_ctx *ctx, int option, int value);
int sph_get_option(const sph_ctx *ctx, int option, int *value);
int sph_world_set_option(sph_world *world, int option, int value);
int sph_world_get_option(const sph_world *world, int option, int *value);

#ifdef __cplusplus
}
#endif
#endif
