"""Worlds: many systems advanced in shared multi-system launches.

Every per-particle expression is evaluated in the same order whether a system
runs alone or inside a world, and every per-system reduction folds only that
system's workgroups in a fixed shape, so a system stepped inside a world must
be bit-identical to the same system stepped alone: state, step size,
iteration counts, error history, measurements and rigid-body state.
"""
import numpy as np
import pytest

from sph_amd import scenes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def native():
    from sph_amd import system
    return system


def assert_same(a, b):
    sa, sb = a.state(), b.state()
    for k in ("x", "v", "a", "V", "p", "m", "fluid_id"):
        assert np.array_equal(sa[k], sb[k]), k
    ma, mb = a.take_measurement(), b.take_measurement()
    assert ma == mb
    assert np.array_equal(a.error_history(), b.error_history())
    for bi in range(len(a.scene.boundaries)):
        assert np.array_equal(a.body_forces(bi), b.body_forces(bi)), bi
        if a.scene.boundaries[bi].rigid is not None:
            ba, bb = a.body_state(bi), b.body_state(bi)
            for k in ba:
                assert np.array_equal(ba[k], bb[k]), (bi, k)


def run_world_vs_solo(native, scs, steps):
    solo = [native.System(sc) for sc in scs]
    mates = [native.System(sc) for sc in scs]
    w = native.World(mates)
    assert w.size() == (len(scs), sum(len(s) for s in mates))
    for _ in range(steps):
        for s in solo:
            s.step_forward_in_time(1)
        w.step(1)
        for a, b in zip(solo, mates):
            assert_same(a, b)
    return solo, mates, w


def test_world_iisph_mixed_sizes_and_bodies(native):
    scs = [scenes.small_iisph_box(side=10, seed=1, jitter=2e-3),
           scenes.small_rigid_scene(),
           scenes.small_iisph_box(side=13, seed=4, dx=0.025, box=0.7, jitter=2e-3),
           scenes.small_iisph_box(side=7, seed=5, jitter=2e-3)]
    scs[2].params.max_time_increment = 4e-4
    scs[3].params.target_density_error = 0.05
    solo, mates, _ = run_world_vs_solo(native, scs, 8)
    # the systems run with different step sizes and iteration counts
    dts = {s.take_measurement()["time_step_size"] for s in solo}
    assert len(dts) > 1


def test_world_scenes_leaving_the_solve_at_different_chunks(native):
    """A pool solved to a 0.02 % target after 40 steps alone (6-10 Jacobi
    iterations per step) beside a small box (3): the box leaves the solve in
    the world's first chunk while the pool runs further chunks; the box's
    iteration count, predicted error and error history stay its own (those of
    the box stepped alone), not the slots of the chunk it last ran in."""
    pool = scenes.pool()
    pool.params.target_density_error = 0.02
    scs = [pool, scenes.small_iisph_box(side=7, seed=5, jitter=2e-3)]
    solo = [native.System(sc) for sc in scs]
    mates = [native.System(sc) for sc in scs]
    solo[0].step_forward_in_time(40)
    mates[0].step_forward_in_time(40)
    w = native.World(mates)
    for _ in range(5):
        for s in solo:
            s.step_forward_in_time(1)
        w.step(1)
        for a, b in zip(solo, mates):
            assert_same(a, b)
        its = [s.take_measurement()["solver_iterations"] for s in solo]
        assert its[0] > its[1] + 2, its


def test_world_with_moving_sample_body(native):
    """A StaticSampleBoundary body that moves (its samples re-posed and
    re-sorted every step) beside a volume-map scene: both bit-identical to
    their solo runs."""
    scs = [scenes.sample_body_scene(), scenes.small_iisph_box(side=9, seed=3, jitter=2e-3)]
    print("world with a moving sample body", flush=True)
    run_world_vs_solo(native, scs, 6)


def test_world_iisphwost(native):
    scs = [scenes.small_iisph_box(side=9, seed=2, solver="IISPHwOST", jitter=2e-3),
           scenes.small_iisph_box(side=12, seed=3, solver="IISPHwOST", jitter=2e-3)]
    run_world_vs_solo(native, scs, 6)


def test_world_sesph_static_samples(native):
    scs = [scenes.dam_break_sesph(side=8, seed=0), scenes.dam_break_sesph(side=11, seed=1)]
    run_world_vs_solo(native, scs, 6)


def verlet_slot(sim):
    return np.asarray(sim.get_checkpoint().fluid.integrator_position_slots[0],
                      dtype=np.float64).reshape(-1, 3)


def test_world_mixed_integrators(native):
    """EulerCromer, ExplicitEuler and Verlet members (IISPH) in one world,
    Verlet's slot set to the positions after System::new: every member and
    Verlet's slot bit-identical to the solo runs."""
    scs = [scenes.small_iisph_box(side=8, seed=1, jitter=2e-3),
           scenes.small_iisph_box(side=9, seed=2, jitter=2e-3),
           scenes.small_iisph_box(side=7, seed=3, jitter=2e-3)]
    scs[1].params.integrator = "ExplicitEuler"
    scs[2].params.integrator = "Verlet"
    print("world with mixed integrators", flush=True)
    solo = [native.System(sc) for sc in scs]
    mates = [native.System(sc) for sc in scs]
    for s in (solo[2], mates[2]):
        s.set_integrator_slots(s.state()["x"])
    w = native.World(mates)
    for _ in range(6):
        for s in solo:
            s.step_forward_in_time(1)
        w.step(1)
        for a, b in zip(solo, mates):
            assert_same(a, b)
        assert np.array_equal(verlet_slot(solo[2]), verlet_slot(mates[2]))


def test_world_sesph_splitting(native):
    """SESPHwSplitting members on static samples and on a volume-map box."""
    scs = [scenes.dam_break_sesph(side=8, seed=0), scenes.small_iisph_box(side=9, seed=2, jitter=2e-3)]
    for sc in scs:
        sc.params.solver = "SESPHwSplitting"
    print("world of SESPHwSplitting members", flush=True)
    run_world_vs_solo(native, scs, 6)


def test_world_compaction(native):
    a = scenes.small_iisph_box(side=8, seed=1, jitter=2e-3)
    a.params.disable_particles_below = float(np.sort(a.x[:, 2])[40]) - 1e-6
    b = scenes.small_iisph_box(side=9, seed=2, jitter=2e-3)
    solo, mates, _ = run_world_vs_solo(native, [a, b], 6)
    assert len(mates[0]) < a.n and len(mates[1]) == b.n


def test_world_member_without_particles(native):
    """A member compacted to no particles steps beside a live one: both stay
    bit-identical to their solo runs."""
    a = scenes.small_iisph_box(side=6, seed=1, jitter=2e-3)
    a.params.disable_particles_below = float(np.max(a.x[:, 2])) + 1.0
    b = scenes.small_iisph_box(side=9, seed=2, jitter=2e-3)
    print("world member without particles", flush=True)
    solo, mates, _ = run_world_vs_solo(native, [a, b], 4)
    assert len(mates[0]) == 0 and len(mates[1]) == b.n


def test_world_attach_detach(native):
    sa = scenes.small_iisph_box(side=9, seed=1, jitter=2e-3)
    sb = scenes.small_iisph_box(side=10, seed=2, jitter=2e-3)
    ref = native.System(sa)
    a, b = native.System(sa), native.System(sb)
    w = native.World([a, b])
    with pytest.raises(native.N.SphError):
        a.step_forward_in_time(1)  # attached systems step through their world
    w.step(2)
    ref.step_forward_in_time(2)
    w.detach(a)
    assert w.size()[0] == 1
    a.step_forward_in_time(2)  # continues alone
    ref.step_forward_in_time(2)
    assert_same(ref, a)
    w.step(1)
    c = native.System(scenes.small_iisph_box(side=6, seed=3, jitter=2e-3))
    w.attach(c)  # joins mid-run
    w.step(2)
    assert w.size() == (2, len(b) + len(c))


def test_world_rejects_mixed_solvers(native):
    a = native.System(scenes.small_iisph_box(side=6, seed=1, jitter=2e-3))
    b = native.System(scenes.small_iisph_box(side=6, seed=1, solver="IISPHwOST", jitter=2e-3))
    w = native.World([a])
    with pytest.raises(native.N.SphError):
        w.attach(b)
    w.step(1)


def test_gradient_cache_modes_bit_identical(native):
    """Pair passes reading cached grad W and recomputing it give the same bits."""
    scs = [scenes.small_rigid_scene(), scenes.small_iisph_box(side=9, seed=2, jitter=2e-3)]
    a = [native.System(sc) for sc in scs]
    b = [native.System(sc) for sc in scs]
    for s in a:
        s.set_cache_gradients(True)
    w = native.World(b)
    w.set_cache_gradients(True)
    for k in range(5):
        for s in a:
            s.step_forward_in_time(1)
        w.step(1)
        if k == 2:  # switch modes mid-run
            for s in a:
                s.set_cache_gradients(False)
            w.set_cache_gradients(False)
    for x, y in zip(a, b):
        assert_same(x, y)
    solo = native.System(scs[1])
    solo.step_forward_in_time(5)
    assert_same(solo, a[1])


@pytest.mark.parametrize("width", [1, 5])
def test_one_sweep_matches_separate_passes(native, width):
    """The one-sweep neighbor list + volume (fixed-width cache) and the count /
    fill / volume passes (slice-width cache) give the same bits, alone and in
    a world; width 5 starts each system with a cache narrower than its
    neighbor counts, so the first sweep runs again wider."""
    from sph_amd import _native as N
    scs = [scenes.small_rigid_scene(), scenes.small_iisph_box(side=9, seed=2, jitter=2e-3),
           scenes.dam_break_sesph(side=8),
           scenes.small_iisph_box(side=9, seed=2, solver="IISPHwOST", jitter=2e-3)]
    a = [native.System(sc, init=False) for sc in scs]        # separate passes
    b = [native.System(sc, init=False) for sc in scs]        # one sweep
    for s in a:
        s.set_option(N.OPT_SWEEP, 0)
    for s in b:
        s.set_option(N.OPT_SWEEP, width)
    for s in a + b:
        N.check(N.lib().sph_system_init(s.handle))
    c = [native.System(sc) for sc in scs[:2]]    # one sweep in a world (one solver)
    w = native.World(c)
    w.set_option(N.OPT_SWEEP, width)
    for k in range(4):
        for s in a + b:
            s.step_forward_in_time(1)
        w.step(1)
        print(f"sweep width {width}: step {k + 1}", flush=True)
    for x, y in zip(a, b):
        assert_same(x, y)
    for x, z in zip(a, c):
        assert_same(x, z)


def test_code_shapes_bit_identical_and_tuner(native):
    """Every code shape of the neighbor passes gives the same bits; the online
    tuner measures all of them on live steps and then keeps one per pass."""
    from sph_amd import _native as N
    sc = scenes.small_iisph_box(side=11, seed=3, solver="IISPHwOST", jitter=2e-3)
    ref = native.System(sc)
    ref.set_option(N.OPT_SHAPE, 0)
    ref.step_forward_in_time(N.SHAPES + 2)
    for shape in range(1, N.SHAPES):
        s = native.System(sc)
        s.set_option(N.OPT_SHAPE, shape)
        s.step_forward_in_time(N.SHAPES + 2)
        assert_same(ref, s)
    tuned = native.System(sc)
    assert tuned.tuning()["state"] == 0
    tuned.step_forward_in_time(N.SHAPES)
    t = tuned.tuning()
    assert t["state"] == 2
    assert all(0 <= t[p] < N.SHAPES for p in N.PAIR_PASSES)
    tuned.step_forward_in_time(2)
    assert_same(ref, tuned)
    # one pass pinned: that pass takes the shape, the others keep theirs, the tuner is off
    mixed = native.System(sc)
    mixed.set_option(N.OPT_SHAPE, 0)
    mixed.set_option(N.OPT_SHAPE_VOLUME + 4, N.SHAPES - 1)
    t = mixed.tuning()
    assert t["jacobi"] == N.SHAPES - 1 and t["pressure_accel"] == 0 and t["steps_per_shape"] == 0
    mixed.step_forward_in_time(N.SHAPES + 2)
    assert_same(ref, mixed)


def test_graph_replay_bit_identical(native):
    """Jacobi chunks replayed from HIP graphs give the eager launches' bits on
    every replay (graphs are captured on a key's second use and then replayed
    every step, with eager work and host waits in between)."""
    from sph_amd import _native as N
    for solver in ("IISPH", "IISPHwOST"):
        sc = scenes.small_iisph_box(side=10, seed=5, solver=solver, jitter=2e-3)
        eager, graph = native.System(sc), native.System(sc)
        for s in (eager, graph):
            s.set_option(N.OPT_SHAPE, 1)      # fixed shapes: keys repeat from step 1
        eager.set_option(N.OPT_GRAPHS, 0)
        assert graph.get_option(N.OPT_GRAPHS) == 1
        for _ in range(8):
            eager.step_forward_in_time(1)
            graph.step_forward_in_time(1)
            assert_same(eager, graph)
    w = native.World([native.System(scenes.small_iisph_box(side=9, seed=k, jitter=2e-3))
                      for k in (1, 2)])
    assert w.get_option(N.OPT_GRAPHS) == 1   # worlds replay graphs too
    solo = [native.System(scenes.small_iisph_box(side=9, seed=k, jitter=2e-3)) for k in (1, 2)]
    for s in solo:
        s.set_option(N.OPT_GRAPHS, 0)
    for _ in range(6):
        w.step(1)
        for s in solo:
            s.step_forward_in_time(1)
    for a, b in zip(solo, w.members):
        assert_same(a, b)


def test_world_graphs_survive_admissions(native):
    """A world that replays its Jacobi chunks from graphs while systems join
    and leave (every join or departure changes the launch layout) keeps each
    system bit-identical to its eager solo run, and its graphs serve every
    layout: their number stays at the few (termination rule, chunk length,
    grid bucket) combinations instead of growing with each layout."""
    from sph_amd import _native as N
    mk = [lambda k=k: scenes.small_iisph_box(side=7 + k, seed=20 + k, jitter=2e-3)
          for k in range(5)]
    world_sys = [native.System(f()) for f in mk]
    solo = [native.System(f()) for f in mk]
    for s in world_sys + solo:
        s.set_option(N.OPT_SHAPE, 1)
    for s in solo:
        s.set_option(N.OPT_GRAPHS, 0)
    w = native.World([world_sys[0], world_sys[1]])
    w.set_option(N.OPT_SHAPE, 1)
    steps = [0] * 5
    plan = [("step", 3), ("attach", 2), ("step", 3), ("attach", 3), ("step", 2),
            ("detach", 0), ("step", 3), ("attach", 4), ("detach", 2), ("step", 4),
            ("detach", 1), ("step", 3)]
    members = [0, 1]
    for op, arg in plan:
        if op == "attach":
            w.attach(world_sys[arg])
            members.append(arg)
        elif op == "detach":
            w.detach(world_sys[arg])
            members.remove(arg)
        else:
            w.step(arg)
            for m in members:
                steps[m] += arg
    counts = w.get_option(N.OPT_GRAPH_COUNT)
    print(f"world graphs after {len(plan)} layout changes: {counts}", flush=True)
    assert 0 < counts <= 12
    for k in range(5):
        solo[k].step_forward_in_time(steps[k]) if steps[k] else None
        assert_same(solo[k], world_sys[k])
    w.close()
