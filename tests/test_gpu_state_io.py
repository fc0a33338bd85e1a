"""Checkpoint resume and recordings on the device.

A system saved after k steps and resumed from the saved file
(SPHSystem::continue_from_checkpoint, sph/mod.rs:362-378) continues bit for
bit as the uninterrupted system when no body moves: the fluid is stored in the
reference's particle order, which is the order the library sorts ties by, and
update() recomputes everything else from it. The continued trajectory is also
checked against the oracle stepping the uninterrupted run. The integrator's
slots are saved but not read back (the reference re-sizes them on resume).

A moving body resumes differently, as in the reference: the saved orientation
is re-normalized (UnitQuaternion::from_quaternion, rigid_body_motion.rs:
273-289), and the saved pending force and torque are restored before the
resuming update() adds that update's fluid and gravity forces on top
(rigid_body_motion.rs:214-222; the accumulators are only cleared by the
body's own step, :183-201), so the first resumed body step integrates twice
the pending force.
"""
import os

import numpy as np
import pytest

import refcpu
from sph_amd import scenes, state_io

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def native():
    from sph_amd import system
    return system


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


def scheme(sc, solver=None, integrator=None):
    sc.params.solver = solver or sc.params.solver
    sc.params.integrator = integrator or sc.params.integrator
    return sc


SCENES = {
    "iisph_box": lambda: scenes.small_iisph_box(side=10, jitter=2e-3),
    "ost_box": lambda: scenes.small_iisph_box(side=10, solver="IISPHwOST", jitter=2e-3),
    "sesph": lambda: scenes.dam_break_sesph(side=8),
    "splitting": lambda: scheme(scenes.dam_break_sesph(side=8), solver="SESPHwSplitting"),
    "explicit_euler": lambda: scheme(scenes.small_iisph_box(side=10, jitter=2e-3),
                                     integrator="ExplicitEuler"),
}


@pytest.mark.parametrize("case", list(SCENES))
@pytest.mark.parametrize("fmt", ["ron", "bincode"])
def test_resume_continues_the_run(native, tmp_path, case, fmt):
    sc = SCENES[case]()
    print(f"resume {case} {fmt}: {sc.n} particles", flush=True)
    a = native.System(sc)
    a.step_forward_in_time(3)
    path = str(tmp_path / f"state.{fmt}")
    state_io.save_checkpoint(a.get_checkpoint(), path, fmt)
    cp = state_io.load_checkpoint(path)
    assert cp.time_steps_propagated == 3
    if sc.params.integrator == "TakePredicted":
        assert len(cp.fluid.integrator_position_slots) == 1
    b = native.System.from_checkpoint(sc, cp)
    ma, mb = a.take_measurement(), b.take_measurement()
    assert mb["steps"] == ma["steps"] and mb["time"] == ma["time"]
    a.step_forward_in_time(3)
    b.step_forward_in_time(3)
    sa, sb = a.state(), b.state()
    for k in ("x", "v", "a", "V", "p", "m"):
        assert np.array_equal(bits(sa[k]), bits(sb[k])), (k, int(np.sum(sa[k] != sb[k])))
    assert np.array_equal(sa["fluid_id"], sb["fluid_id"])
    ma, mb = a.take_measurement(), b.take_measurement()
    assert ma["steps"] == mb["steps"] == 6
    assert ma["solver_iterations"] == mb["solver_iterations"]
    # the uninterrupted run itself follows the oracle
    o = refcpu.OracleSystem(sc)
    o.step(6)
    so = o.state()
    for k in ("x", "v", "V", "p"):
        den = float(np.max(np.abs(so[k]))) or 1.0
        assert float(np.max(np.abs(sb[k] - so[k]))) / den <= 1e-5, k


@pytest.mark.parametrize("fmt", ["ron", "bincode"])
def test_resume_resets_the_verlet_slot(native, tmp_path, fmt):
    """Verlet's position slot (the position one step back, verlet.rs:10-34)
    is saved with the fluid (fluid.rs:352-381), but both of the reference's
    resume paths re-size the slots before the first step (sph/mod.rs:368-373,
    404-409), so the resumed run starts again from the origin slot. It then
    follows the oracle started from the checkpoint's fluid bit for bit (under
    SESPH: the origin slot throws every particle to about twice its position,
    and IISPH's solve, as the reference's, then does not converge)."""
    import dataclasses
    sc = scenes.dam_break_sesph(side=6)
    sc.params.integrator = "Verlet"
    print(f"resume verlet {fmt}", flush=True)
    a = native.System(sc)
    a.set_integrator_slots(sc.x)
    a.step_forward_in_time(3)
    path = str(tmp_path / f"state.{fmt}")
    state_io.save_checkpoint(a.get_checkpoint(), path, fmt)
    cp = state_io.load_checkpoint(path)
    assert len(cp.fluid.integrator_position_slots) == 1
    assert len(cp.fluid.integrator_velocity_slots) == 0
    saved = np.asarray(cp.fluid.integrator_position_slots[0], dtype=np.float64).reshape(-1, 3)
    assert np.count_nonzero(saved) > 0
    b = native.System.from_checkpoint(sc, cp)
    xp = b.get_checkpoint().fluid.integrator_position_slots[0]
    assert not np.any(np.asarray(xp, dtype=np.float64))
    f = cp.fluid
    o = refcpu.OracleSystem(dataclasses.replace(
        sc, x=np.asarray(f.position, dtype=np.float64).reshape(-1, 3),
        v=np.asarray(f.velocity, dtype=np.float64).reshape(-1, 3),
        m=np.asarray(f.mass, dtype=np.float64).ravel(),
        fluid_id=np.asarray(f.fluid_id, dtype=np.uint32).ravel()))
    for k in range(2):
        o.step(1)
        b.step_forward_in_time(1)
        so, sb = o.state(), b.state()
        for f_ in ("x", "v", "a", "V", "p"):
            assert np.array_equal(bits(sb[f_]), bits(so[f_])), (k, f_)


@pytest.mark.parametrize("case", ["volume_map_body", "sample_body"])
def test_resume_restores_rigid_state(native, case):
    """A moving body resumes its rigid state; samples on a moving
    StaticSampleBoundary body follow the restored pose
    (static_sample_boundary.rs:560-580) with the volumes of their initial
    pose."""
    sc = scenes.small_rigid_scene() if case == "volume_map_body" else scenes.sample_body_scene()
    print(f"resume rigid state: {case}", flush=True)
    a = native.System(sc)
    a.step_forward_in_time(4)
    cp = a.get_checkpoint()
    dyn = [b for b, d in enumerate(cp.dynamic_states) if d is not None]
    assert dyn, "the scene has a moving body"
    assert all((cp.dynamic_states[b] is None) == (getattr(d, "rigid", None) is None)
               for b, d in enumerate(sc.boundaries))
    b = native.System.from_checkpoint(sc, cp)
    for k in dyn:
        ra, rb = a.body_state(k), b.body_state(k)
        for f in ("com", "lin_vel", "ang_mom"):
            assert np.array_equal(ra[f], rb[f]), f
        assert np.max(np.abs(ra["q"] - rb["q"])) <= 4e-16
        # pending force and torque: the saved ones plus the resuming update's
        # own, which repeat the saved ones up to the re-normalized orientation
        for f in ("force", "torque"):
            scale = 1.0 + float(np.max(np.abs(ra[f])))
            assert np.max(np.abs(rb[f] - 2.0 * ra[f])) <= 1e-9 * scale, f
        fa, fb = a.body_forces(k), b.body_forces(k)
        assert np.max(np.abs(fb - fa)) <= 1e-9 * (1.0 + float(np.max(np.abs(fa))))
    # a checkpoint with the wrong number of boundaries leaves bodies untouched
    bad = state_io.SystemCheckpoint(cp.time_steps_propagated, cp.current_time, cp.fluid,
                                    cp.dynamic_states + [None])
    c = native.System.from_checkpoint(sc, bad)
    init = native.System(sc, init=True)
    for k in dyn:
        assert np.array_equal(c.body_state(k)["com"], init.body_state(k)["com"])


def test_recording_of_a_run(native, tmp_path):
    sc = scenes.small_iisph_box(side=10, jitter=2e-3)
    sim = native.System(sc)
    path = str(tmp_path / "run.bin")
    rec = state_io.Recording(path, state_io.SimulationParameters(
        float(sc.params.rest_density_grid_spacing), (0.0, 0.0, 5.0), 100, True, True, False))
    u = np.linspace(0.0, 0.6, 8, dtype=np.float32)
    gx, gy = np.meshgrid(u, u)
    plane = np.stack([gx.ravel(), gy.ravel(), np.full(gx.size, 0.1, np.float32)], axis=1)
    for k in range(3):
        sim.step_forward_in_time(1)
        rec.append_time_step(sim.time_step_info(
            {"kind": "Samples", "coloring": {"kind": "QuantityGraded", "quantity": "Density"}}))
    # the same step number again replaces that step's record
    rec.append_time_step(sim.time_step_info(
        {"kind": "SensorPlane", "quantity": "Pressure",
         "planes": [{"positions": plane, "rows": 8, "cols": 8}]}))
    sp, steps = state_io.read_recording(path)
    assert [s["time_step_number"] for s in steps] == [1, 2, 3]
    last = steps[-1]
    assert last["fluid"]["kind"] == "SensorPlane"
    got = last["fluid"]["planes"][0]["data"]
    assert np.array_equal(got, sim.get_quantity_at_positions("Pressure", plane))
    st = sim.state()
    m = sim.take_measurement()
    assert last["measurement"]["solver_iterations"] == m["solver_iterations"]
    # SolverMeasurementInfo of IISPH (iisph.rs:98-106): no stiffness
    mm = last["measurement"]
    assert mm["stiffness"] == 0.0 and mm["time_step_wall_clock_time"] > 0.0
    assert mm["target_density_error"] == sc.params.target_density_error
    assert mm["relaxation_factor"] == sc.params.relaxation_factor
    assert steps[1]["fluid"]["positions"].shape == (sc.n, 3)
    dens = steps[1]["fluid"]["coloring"]["data"]
    assert dens.dtype == np.float32 and np.all(dens > 0)
    assert os.path.getsize(path) > 0 and st["x"].shape == (sc.n, 3)


def test_trait_accessors(native):
    """The SPHSystem trait's accessors (sph/mod.rs:40-71, 107-238): time, step
    count, fluid ids and f32 positions in the reference's particle order, the
    fluid and boundary checkpoints, per-particle quantities."""
    sc = scenes.small_rigid_scene()
    sim = native.System(sc)
    sim.step_forward_in_time(2)
    st, m = sim.state(), sim.take_measurement()
    assert sim.time() == m["time"] > 0 and sim.time_steps_propagated() == 2
    assert np.array_equal(sim.get_fluid_ids(), st["fluid_id"])
    pos = sim.get_fluid_pos()
    assert pos.dtype == np.float32 and np.array_equal(pos, st["x"].astype(np.float32))
    f = sim.get_fluid_checkpoint()
    assert np.array_equal(np.asarray(f.position).reshape(-1, 3), st["x"])
    b = sim.get_boundary_checkpoint()
    assert len(b) == len(sc.boundaries) and any(e is not None for e in b)
    # continue_from_checkpoint on a live system: the saved state and counters
    cp = sim.get_checkpoint()
    sim.step_forward_in_time(3)
    sim.continue_from_checkpoint(cp)
    assert sim.time_steps_propagated() == 2
    assert np.array_equal(sim.state()["x"], st["x"])
    rho = sim.get_quantity_of_fluid_samples("Density")
    assert np.array_equal(rho, (st["m"] / st["V"]).astype(np.float32))
    assert np.array_equal(sim.get_quantity_of_fluid_samples(2), rho)


@pytest.mark.parametrize("solver", ["SESPH", "SESPHwSplitting"])
def test_recorded_measurement_of_state_equation_solvers(native, solver):
    """SolverMeasurementInfo of SESPH and SESPHwSplitting (sesph.rs:69-74,
    sesph_with_splitting.rs:76-81): the stiffness, and zero iterations, target
    error, relaxation factor and predicted error."""
    sc = scenes.dam_break_sesph(side=6)
    sc.params.solver = solver
    sim = native.System(sc)
    sim.step_forward_in_time(2)
    m = sim.time_step_info()["measurement"]
    assert m["stiffness"] == sc.params.stiffness > 0
    for k in ("target_density_error", "relaxation_factor", "predicted_density_error",
              "solver_iterations"):
        assert m[k] == 0, k


def test_fixed_step_time_is_offset_plus_steps_times_dt(native, tmp_path):
    """Without CFL the reference reports time_offset + (steps - offset_steps) *
    time_increment (sph/mod.rs:107-116), not a running sum: after 10 steps of
    5e-4 the time is exactly 10 * 5e-4; a system built from a checkpoint takes
    the saved time as its offset (sph/mod.rs:408-417). The oracle agrees."""
    sc = scenes.dam_break_sesph(side=8)
    dt = sc.params.time_increment
    a = native.System(sc)
    a.step_forward_in_time(10)
    assert a.take_measurement()["time"] == 10 * dt
    o = refcpu.OracleSystem(sc)
    o.step(10)
    assert o.measure()["time"] == 10 * dt
    path = str(tmp_path / "state.ron")
    state_io.save_checkpoint(a.get_checkpoint(), path, "ron")
    b = native.System.from_checkpoint(sc, state_io.load_checkpoint(path))
    b.step_forward_in_time(7)
    t0 = state_io.load_checkpoint(path).current_time
    assert b.take_measurement()["time"] == t0 + 7 * dt
