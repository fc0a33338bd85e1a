"""GPU parity: the HIP path (through the C ABI) against the CPU oracle.

Bars (from the hot-path contract): neighbor sets and cell sort bit-exact;
per-step density (volume), pressure, acceleration and body force/torque within
1e-5 relative (norm-wise per field); IISPH iteration counts equal.
The HIP path computes in binary64 with the reference's evaluation order, so
the observed differences are at round-off level; the stated tolerance is the
contract, the tighter checks below document what is achieved.
"""
import numpy as np
import pytest

from sph_amd import scenes

pytestmark = pytest.mark.gpu

REL_TOL = 1e-5


def rel_err(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.max(np.abs(b)) if b.size else 0.0
    if den == 0.0:
        return float(np.max(np.abs(a))) if a.size else 0.0
    return float(np.max(np.abs(a - b)) / den)


@pytest.fixture(scope="module")
def native():
    from sph_amd import system
    return system


@pytest.fixture(scope="module")
def refcpu():
    import refcpu
    return refcpu


def compare_states(gs, os_, fields=("x", "v", "a", "V", "p"), tol=REL_TOL):
    errs = {}
    for k in fields:
        errs[k] = rel_err(gs[k], os_[k])
        assert errs[k] <= tol, (k, errs[k])
    return errs


# ---------------------------------------------------------------- neighbors
@pytest.mark.parametrize("case", ["random", "lattice", "boundary", "wide_range", "negative"])
def test_neighbors_bit_exact(native, refcpu, case):
    rng = np.random.default_rng(7)
    if case == "random":
        pos = rng.uniform(-0.3, 0.3, size=(3000, 3))
        smp = pos
        cs = rng_ = 0.04
    elif case == "lattice":
        sc = scenes.dam_break_sesph()
        pos = smp = sc.x
        cs = rng_ = sc.params.kernel_support_radius
    elif case == "boundary":
        sc = scenes.dam_break_sesph()
        pos, smp = sc.x, sc.boundaries[0].positions
        cs = rng_ = sc.params.kernel_support_radius
    elif case == "wide_range":  # range 2x the cell size: 5^3 stencil
        pos = rng.uniform(0, 1, size=(2000, 3))
        smp = rng.uniform(0, 1, size=(2500, 3))
        cs, rng_ = 0.05, 0.1
    else:  # cells on both sides of zero exercise both map_int_to_uint branches
        pos = rng.uniform(-0.2, 0.2, size=(2000, 3))
        smp = rng.uniform(-0.2, 0.2, size=(2000, 3))
        cs = rng_ = 0.05
    sh = native.SpatialHashing(cs)
    nl = sh.find_samples(rng_, pos, smp)
    off, idx = refcpu.find_samples(cs, rng_, pos, smp)
    assert np.array_equal(nl.offsets, off)
    assert np.array_equal(nl.indices, idx)  # same sets, same order


def test_neighbors_exact_range_excluded(native):
    # distance exactly equal to the range is excluded (strict <)
    pos = np.array([[0.0, 0.0, 0.0], [2.0, 0.0, 0.0], [1.0, 0.0, 0.0]])
    nl = native.SpatialHashing(1.0).find_samples(2.0, pos, pos)
    assert list(nl.get_neighbors(0)) == [0, 2]
    assert 0 not in list(nl.get_neighbors(1))


def test_neighbors_empty_inputs(native):
    sh = native.SpatialHashing(0.1)
    nl = sh.find_samples(0.1, np.zeros((0, 3)), np.zeros((5, 3)))
    assert nl.num_samples() == 0 and len(nl.indices) == 0
    nl = sh.find_samples(0.1, np.zeros((4, 3)), np.zeros((0, 3)))
    assert nl.num_samples() == 4 and len(nl.indices) == 0


def test_system_neighbor_list_matches_oracle(native, refcpu):
    sc = scenes.small_iisph_box()
    o = refcpu.OracleSystem(sc)
    off, idx = o.fluid_neighbors()
    sh = native.SpatialHashing(sc.params.kernel_support_radius)
    st = o.state()
    nl = sh.find_samples(sc.params.kernel_support_radius, st["x"], st["x"])
    assert np.array_equal(nl.offsets, off) and np.array_equal(nl.indices, idx)


# ---------------------------------------------------------------- per-trait
def test_volume_and_boundary_samples(native, refcpu):
    sc = scenes.small_iisph_box(side=10, jitter=2e-3)
    o = refcpu.OracleSystem(sc)
    st = o.state()
    g = native.System(sc, init=False)
    V = g.get_volume(st["x"])
    assert rel_err(V, st["V"]) <= 1e-12
    samples, flags = g.find_boundary_samples(0, st["x"])
    for i in range(sc.n):
        ref = o.boundary_samples(0, i, 0)
        assert int(flags[i]) == len(ref)
        if len(ref):
            assert np.allclose(samples[i], ref[0], rtol=0, atol=1e-14)


def test_nonpressure_and_pressure_solve(native, refcpu):
    for solver in ("SESPH", "IISPH", "IISPHwOST", "SESPHwSplitting"):
        sc = scenes.small_iisph_box(side=10, solver=solver, jitter=2e-3)
        if solver in ("SESPH", "SESPHwSplitting"):
            sc.params.integrator = "EulerCromer"
        o = refcpu.OracleSystem(sc)
        o.step(3)
        st = o.state()
        m = o.measure()
        g = native.System(sc, init=False)
        V = g.get_volume(st["x"])
        a_np, _ = g.add_non_pressure_accelerations(st["x"], st["v"], st["m"], V)
        res = g.solve_and_add_acceleration(st["x"], st["v"], st["m"], V, a_np,
                                           m["time_step_size"])
        assert rel_err(res["p"], st["p"]) <= REL_TOL
        assert rel_err(res["a"], st["a"]) <= REL_TOL
        if solver != "SESPH":
            assert res["info"]["solver_iterations"] == m["solver_iterations"]


# ---------------------------------------------------------------- stepping
def run_pair(native, refcpu, sc, steps):
    o = refcpu.OracleSystem(sc)
    g = native.System(sc)
    errs = [compare_states(g.state(), o.state())]
    mo, mg = o.measure(), g.take_measurement()
    assert mg["solver_iterations"] == mo["solver_iterations"]
    for _ in range(steps):
        o.step(1)
        g.step_forward_in_time(1)
        mo, mg = o.measure(), g.take_measurement()
        assert mg["solver_iterations"] == mo["solver_iterations"]
        assert mg["eqs1_iterations"] == mo["eqs1_iterations"]
        assert abs(mg["time_step_size"] - mo["time_step_size"]) <= 1e-12 * mo["time_step_size"]
        errs.append(compare_states(g.state(), o.state()))
        assert abs(mg["density_error"] - mo["density_error"]) <= 1e-6 + 1e-5 * abs(mo["density_error"])
    return o, g, errs


def test_sesph_dam_break_steps(native, refcpu):
    run_pair(native, refcpu, scenes.dam_break_sesph(), 10)


def test_iisph_box_steps(native, refcpu):
    o, g, _ = run_pair(native, refcpu, scenes.small_iisph_box(side=14, jitter=2e-3), 15)
    he, hg = o.error_history(), g.error_history()
    assert len(he) == len(hg)
    assert np.allclose(hg, he, rtol=1e-9, atol=1e-15)


def test_iisphwost_box_steps(native, refcpu):
    run_pair(native, refcpu, scenes.small_iisph_box(side=12, solver="IISPHwOST", jitter=2e-3), 12)


def test_rigid_coupling_steps(native, refcpu):
    sc = scenes.small_rigid_scene()
    o, g, _ = run_pair(native, refcpu, sc, 25)
    for b in range(len(sc.boundaries)):
        fo, fg = o.body_forces(b), g.body_forces(b)
        assert rel_err(fg, fo) <= REL_TOL, (b, fg, fo)
        if sc.boundaries[b].rigid is None:
            continue
        so, sg = o.body_state(b), g.body_state(b)
        for k in ("com", "q", "lin_vel", "ang_mom"):
            assert rel_err(sg[k], so[k]) <= REL_TOL, (b, k)


def test_compaction_below_threshold(native, refcpu):
    sc = scenes.small_iisph_box(side=8, jitter=2e-3)
    # particles of the lowest layer fall below the threshold after a few steps
    sc.params.disable_particles_below = float(np.sort(sc.x[:, 2])[40]) - 1e-6
    o, g, _ = run_pair(native, refcpu, sc, 6)
    assert len(g) == len(o) < sc.n
    assert np.array_equal(g.state()["fluid_id"], o.state()["fluid_id"])


def test_compaction_keeps_the_verlet_slot(native, refcpu):
    """Fluid::swap moves the integrator's slots with their particle and
    drop_inactive truncates them (fluid.rs:159-210): IISPH + Verlet, the slot
    at the positions after System::new, particles of the lowest layer dropped
    as they fall below disable_particles_below; state and slot bit-identical
    to the oracle's."""
    sc = with_scheme(scenes.small_iisph_box(side=8, jitter=2e-3), integrator="Verlet")
    sc.params.disable_particles_below = float(np.sort(sc.x[:, 2])[40]) - 1e-6
    print("compaction with Verlet's slot", flush=True)
    o, g = refcpu.OracleSystem(sc), native.System(sc)
    n0 = len(g)
    o.set_field("slot_x", o.state()["x"])
    g.set_integrator_slots(g.state()["x"])
    for k in range(8):
        o.step(1)
        g.step_forward_in_time(1)
        so, sg = o.state(), g.state()
        for f in ("x", "v", "a", "V", "p", "fluid_id"):
            assert np.array_equal(sg[f], so[f]), (k, f)
        slot = np.asarray(g.get_checkpoint().fluid.integrator_position_slots[0],
                          dtype=np.float64).reshape(-1, 3)
        assert np.array_equal(slot, o.field("slot_x")), k
    assert len(g) == len(o) < n0


@pytest.mark.parametrize("solver", ["IISPH", "IISPHwOST", "SESPH"])
def test_compaction_to_no_particles(native, refcpu, solver):
    """Every particle below disable_particles_below: System::new's update
    disables them all (fluid.rs:159-210) and the empty system keeps stepping,
    as the oracle's does (no iterations, zero errors, time advancing)."""
    sc = scenes.small_iisph_box(side=6, jitter=2e-3, solver=solver)
    sc.params.disable_particles_below = float(np.max(sc.x[:, 2])) + 1.0
    print("compaction to no particles:", solver, flush=True)
    o, g, _ = run_pair(native, refcpu, sc, 3)
    assert len(g) == len(o) == 0
    mg, mo = g.take_measurement(), o.measure()
    assert mg["num_particles"] == mo["num_particles"] == 0
    for k in ("density_error", "solver_iterations", "time_step_size", "time"):
        assert mg[k] == mo[k], (k, mg[k], mo[k])


def test_clone_is_independent(native, refcpu):
    sc = scenes.small_iisph_box(side=10, jitter=2e-3)
    g = native.System(sc)
    g.step_forward_in_time(2)
    c = g.clone()
    g.step_forward_in_time(3)
    c.step_forward_in_time(3)
    sg, scl = g.state(), c.state()
    for k in ("x", "v", "a", "V", "p"):
        assert rel_err(scl[k], sg[k]) <= 1e-12


def test_block_drop_full_size_steps(native, refcpu):
    """The 125k-particle IISPH block drop: a few steps against the oracle."""
    run_pair(native, refcpu, scenes.block_drop_iisph(), 3)


# ---------------------------------------------------------------- bit exactness
@pytest.mark.parametrize("case", ["iisph_box", "ost_box", "rigid", "sesph"])
def test_trajectories_bit_identical(native, refcpu, case):
    """Beyond the 1e-5 contract: the HIP path evaluates every per-particle
    expression in the reference's order in binary64 without contraction, so
    positions, velocities, accelerations, volumes and pressures equal the
    oracle's bit for bit over several steps, and the iteration counts match;
    only the Jacobi error's cross-particle sum differs in grouping."""
    sc = {"iisph_box": lambda: scenes.small_iisph_box(side=12, jitter=2e-3),
          "ost_box": lambda: scenes.small_iisph_box(side=12, solver="IISPHwOST", jitter=2e-3),
          "rigid": scenes.small_rigid_scene,
          "sesph": lambda: scenes.dam_break_sesph(side=10)}[case]()
    o, g = refcpu.OracleSystem(sc), native.System(sc)
    for _ in range(8):
        o.step(1)
        g.step_forward_in_time(1)
        so, sg = o.state(), g.state()
        for k in ("x", "v", "a", "V", "p"):
            assert np.array_equal(sg[k], so[k]), k
        assert g.take_measurement()["solver_iterations"] == o.measure()["solver_iterations"]
    he, hg = o.error_history(), g.error_history()
    assert len(he) == len(hg) and np.allclose(hg, he, rtol=1e-12, atol=1e-18)


def with_twins(sc, dup):
    """The scene with copies of particles `dup` at exactly their positions."""
    sc.x = np.concatenate([sc.x, sc.x[dup]])
    sc.v = np.concatenate([sc.v, sc.v[dup]])
    sc.m = np.concatenate([sc.m, sc.m[dup]])
    sc.fluid_id = np.concatenate([sc.fluid_id, sc.fluid_id[dup]])
    return sc


TWINS = np.array([3, 77, 200, 200, 411])  # particle 200 three times over


def test_coincident_particles_bit_identical(native, refcpu):
    """Distinct particles at exactly the same position (distance 0), SESPH:
    grad W(0) is the zero vector (kernel/mod.rs:37-45) and the viscosity
    denominator's 0.01 dx^2 term (non_pressure_accelerations.rs:106-185) keeps
    every sum finite; both sides list the twins as neighbors and agree bit for
    bit."""
    sc = with_twins(scenes.dam_break_sesph(side=8), TWINS)
    print(f"coincident particles: {sc.n} particles, {len(TWINS)} twins", flush=True)
    o, g = refcpu.OracleSystem(sc), native.System(sc)
    for _ in range(4):
        o.step(1)
        g.step_forward_in_time(1)
        so, sg = o.state(), g.state()
        for k in ("x", "v", "a", "V", "p"):
            assert np.all(np.isfinite(sg[k])), k
            assert np.array_equal(sg[k], so[k]), k


def test_coincident_particles_stall_iisph_loudly(native):
    """Twins never separate under IISPH: pressure gradients between them are
    zero, and the Jacobi error stays above the target. The reference's loop
    has no cap (iisph.rs:387-390) and would not return; the library stops at
    its 100,000-iteration cap with SPH_ERR_UNSUPPORTED instead of hanging."""
    from sph_amd import _native as N
    sc = with_twins(scenes.small_iisph_box(side=8, jitter=2e-3), TWINS)
    print(f"coincident particles under IISPH: {sc.n} particles", flush=True)
    with pytest.raises(N.SphError, match="did not converge"):
        native.System(sc)  # System::new already runs an update with a solve


FULL_SIZE = {
    "ball_IISPHwOST": lambda: scenes.two_fluid_ball(solver="IISPHwOST"),
    "ball_IISPH": lambda: scenes.two_fluid_ball(solver="IISPH"),
    "block_drop": scenes.block_drop_iisph,
    "rigid_coupling": scenes.rigid_coupling,
}


@pytest.mark.parametrize("case", list(FULL_SIZE))
def test_headline_scene_bit_identical_at_full_size(native, refcpu, case):
    """The measured workloads at their full sizes: the 1,000,000-particle
    two-fluid ball (both rest densities, the sphere volume map, CFL dt) with
    either solver, the 125,000-particle block drop and the 492,960-particle
    rigid coupling (three two-way coupled bodies, contact on) equal the oracle
    bit for bit over two steps, with equal iteration counts and rigid states.
    The GPU path runs the sorted, paired pair cache, the tuner's first shapes
    and graph replay; the oracle runs its own hash grid on the host's
    threads."""
    sc = FULL_SIZE[case]()
    print(f"full-size {case}: {sc.n} particles", flush=True)
    o, g = refcpu.OracleSystem(sc), native.System(sc)
    for step in range(2):
        o.step(1)
        g.step_forward_in_time(1)
        so, sg = o.state(), g.state()
        for k in ("x", "v", "a", "V", "p"):
            assert np.array_equal(sg[k], so[k]), (step, k)
        mo, mg = o.measure(), g.take_measurement()
        assert mg["solver_iterations"] == mo["solver_iterations"]
        assert mg["eqs1_iterations"] == mo["eqs1_iterations"]
        for b, bd in enumerate(sc.boundaries):
            if getattr(bd, "rigid", None) is not None:
                gb, ob = g.body_state(b), o.body_state(b)
                for k in ("com", "q", "lin_vel", "ang_mom"):
                    assert np.array_equal(gb[k], ob[k]), (step, b, k)
        print(f"  step {step}: iterations {mo['eqs1_iterations']} + {mo['solver_iterations']}",
              flush=True)


def test_per_trait_calls_bit_identical(native, refcpu):
    for solver in ("SESPH", "IISPH", "IISPHwOST", "SESPHwSplitting"):
        sc = scenes.small_iisph_box(side=10, solver=solver, jitter=2e-3)
        if solver in ("SESPH", "SESPHwSplitting"):
            sc.params.integrator = "EulerCromer"
        o = refcpu.OracleSystem(sc)
        o.step(3)
        st, m = o.state(), o.measure()
        g = native.System(sc, init=False)
        V = g.get_volume(st["x"])
        assert np.array_equal(V, st["V"])
        a_np, _ = g.add_non_pressure_accelerations(st["x"], st["v"], st["m"], V)
        res = g.solve_and_add_acceleration(st["x"], st["v"], st["m"], V, a_np,
                                           m["time_step_size"])
        assert np.array_equal(res["p"], st["p"]) and np.array_equal(res["a"], st["a"])


@pytest.mark.parametrize("solver", ["IISPH", "SESPH", "SESPHwSplitting"])
def test_moving_static_sample_body(native, refcpu, solver):
    """A StaticSampleBoundary body (static_sample_boundary.rs:79-130): samples
    on a cube's surface follow its rigid motion (update_positions_and_velocities
    after every body step), carry its velocity into the viscosity and source
    terms, keep the pseudo volumes of their initial pose, and feed the
    fluid's reactions back into the body. The body's force total is a sum the
    oracle and the GPU group differently, so they agree to rounding: every step
    compares the iteration counts, the step size, the body force and the
    rigid state, every 40 steps the fluid. A clone carries on bit for bit."""
    sc = scenes.sample_body_scene(solver=solver)
    print(f"moving sample body: {solver}, {sc.n} particles", flush=True)
    o, g = refcpu.OracleSystem(sc), native.System(sc)
    worst = {"force": 0.0, "state": 0.0, "fluid": 0.0}
    for k in range(1, 121):
        o.step(1)
        g.step_forward_in_time(1)
        mo, mg = o.measure(), g.take_measurement()
        assert mg["solver_iterations"] == mo["solver_iterations"], (k, mg, mo)
        assert mg["time_step_size"] == pytest.approx(mo["time_step_size"], rel=1e-9), k
        worst["force"] = max(worst["force"], rel_err(g.body_forces(1), o.body_forces(1)))
        so, sg = o.body_state(1), g.body_state(1)
        for f in ("com", "q", "lin_vel", "ang_mom"):
            worst["state"] = max(worst["state"], rel_err(sg[f], so[f]))
        if k % 40 == 0:
            a, b = o.state(), g.state()
            for f in ("x", "v", "a", "V", "p"):
                worst["fluid"] = max(worst["fluid"], rel_err(b[f], a[f]))
    print("  largest relative differences", worst, flush=True)
    assert all(v < 1e-5 for v in worst.values()), worst
    c = g.clone()
    g.step_forward_in_time(5)
    c.step_forward_in_time(5)
    sg, scl = g.state(), c.state()
    for f in ("x", "v", "a", "V", "p"):
        assert np.array_equal(scl[f], sg[f]), f
    assert np.array_equal(c.body_state(1)["com"], g.body_state(1)["com"])


def test_moving_static_sample_body_refused_by_iisphwost(native):
    """IISPHwOST does not support dynamic boundaries (SUPPORTS_DYNAMIC_BOUNDARIES,
    setup/mod.rs:242-244)."""
    from sph_amd import _native as N
    sc = scenes.sample_body_scene(solver="IISPH")
    sc.params.solver, sc.params.integrator = "IISPHwOST", "TakePredicted"
    with pytest.raises(N.SphError, match="IISPHwOST does not support dynamic boundaries"):
        native.System(sc)



# ------------------------------------------- other integrators and solvers
def compacting(sc):
    sc.params.disable_particles_below = float(np.sort(sc.x[:, 2])[40]) - 1e-6
    return sc


def with_scheme(sc, solver=None, integrator=None):
    if solver:
        sc.params.solver = solver
    if integrator:
        sc.params.integrator = integrator
    return sc


SCHEMES = {
    # SESPH + ExplicitEuler on the dam break's static samples
    "sesph_explicit_euler": lambda: with_scheme(scenes.dam_break_sesph(side=8),
                                                integrator="ExplicitEuler"),
    # IISPH + Verlet from a position slot equal to the positions (at rest)
    "iisph_verlet": lambda: with_scheme(scenes.small_iisph_box(side=10, jitter=2e-3),
                                        integrator="Verlet"),
    # Verlet from the slot Fluid::resize_slots leaves (the origin)
    "sesph_verlet_origin_slot": lambda: with_scheme(scenes.dam_break_sesph(side=6),
                                                    integrator="Verlet"),
    # SESPHwSplitting on a volume-map box, on static samples, and with a
    # two-way coupled body
    "splitting_box": lambda: with_scheme(scenes.small_iisph_box(side=10, jitter=2e-3),
                                         solver="SESPHwSplitting"),
    "splitting_static": lambda: with_scheme(scenes.dam_break_sesph(side=8),
                                            solver="SESPHwSplitting"),
    "splitting_rigid": lambda: with_scheme(scenes.small_rigid_scene(), solver="SESPHwSplitting"),
    # SESPHwSplitting (its pair cache keeps the self pair) while compaction
    # drops the lowest layer
    "splitting_compaction": lambda: compacting(with_scheme(
        scenes.small_iisph_box(side=8, jitter=2e-3), solver="SESPHwSplitting")),
}


@pytest.mark.parametrize("case", list(SCHEMES))
def test_other_schemes_bit_identical(native, refcpu, case):
    """ExplicitEuler and Verlet (integration_schemes/explicit_euler.rs,
    verlet.rs) and SESPHwSplitting (pressure_solver/sesph_with_splitting.rs):
    positions, velocities, accelerations, volumes and pressures equal the
    oracle's bit for bit while no moving body couples (the body's force total
    is grouped differently on the two sides), within 1e-9 otherwise."""
    sc = SCHEMES[case]()
    print(f"scheme {case}: {sc.n} particles", flush=True)
    o, g = refcpu.OracleSystem(sc), native.System(sc)
    if case == "iisph_verlet":
        o.set_field("slot_x", sc.x)  # both systems in the reference's particle order
        g.set_integrator_slots(sc.x)
    exact = case != "splitting_rigid"
    for k in range(3 if case == "sesph_verlet_origin_slot" else 8):
        o.step(1)
        g.step_forward_in_time(1)
        so, sg = o.state(), g.state()
        for f in ("x", "v", "a", "V", "p"):
            if exact:
                assert np.array_equal(sg[f], so[f]), (k, f)
            else:
                assert rel_err(sg[f], so[f]) <= 1e-9, (k, f)
        mg, mo = g.take_measurement(), o.measure()
        assert mg["solver_iterations"] == mo["solver_iterations"]
        # the average (a sum over particles, grouped differently) minus 1
        assert abs(mg["density_error"] - mo["density_error"]) <= 1e-12 + 1e-9 * abs(mo["density_error"])
    if "verlet" in case:
        assert g.integrator_slot_counts() == (1, 0)
