"""Rigid-rigid contact on the device against the oracle
(VolumeMapBoundary::step_forward_in_time, volume_map_boundary.rs:56-1102,
1440-1453).

The contact cells come from k_contact_cells (one workgroup per candidate
sub-cell; the cell sums are sequential in quadrature-point order) and the
NNCG solve runs on the host over the cells; both are bit-identical to the
oracle: every field of every active cell record (volume, centre,
sensitivities, diagonal elements, p, lambda), the iteration count, the rigid
bodies and the fluid.
"""
import numpy as np
import pytest

import refcpu
from sph_amd import scenes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def native():
    from sph_amd import system
    return system


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


def compare(g, o, sc):
    assert g.contact_stats() == o.contact_stats()
    np.testing.assert_array_equal(bits(g.contact_cells()), bits(o.contact_cells()))
    for b, bd in enumerate(sc.boundaries):
        if bd.rigid is None:
            continue
        gs, os_ = g.body_state(b), o.body_state(b)
        for k in ("com", "q", "lin_vel", "ang_mom", "force", "torque", "ang_vel"):
            np.testing.assert_array_equal(bits(gs[k]), bits(os_[k]), err_msg=f"body {b} {k}")
    gs, os_ = g.state(), o.state()
    for k in ("x", "v", "a", "V", "p"):
        np.testing.assert_array_equal(bits(gs[k]), bits(os_[k]), err_msg=k)


CASES = {
    "stack": lambda: scenes.contact_stack(),
    "stack_flat_cube": lambda: scenes.contact_stack(tilt=0.0, sphere=False, depth=0.004),
    "halfspace": lambda: scenes.halfspace_contact(0.006),
}


@pytest.mark.parametrize("case", list(CASES))
def test_contact_steps_match_oracle(native, case):
    sc = CASES[case]()
    nb = len(sc.boundaries)
    print(f"contact {case}: {sc.n} particles, {nb} boundaries", flush=True)
    g, o = native.System(sc), refcpu.OracleSystem(sc)
    seen = 0
    for step in range(3):
        g.step_forward_in_time(1)
        o.step(1)
        print(f"  step {step}: cells/iterations {o.contact_stats()}", flush=True)
        compare(g, o, sc)
        seen += o.contact_stats()[0]
    assert seen > 0


def test_contact_cells_beyond_the_first_capacity(native):
    """More active cells (5184) than the buffers first hold (4096): the
    counting pass sizes them and the cells are computed again."""
    sc = scenes.halfspace_contact(0.005, dx=0.01, half=0.36, floor_half=0.4)
    print(f"contact large: {sc.n} particles", flush=True)
    g, o = native.System(sc), refcpu.OracleSystem(sc)
    g.step_forward_in_time(1)
    o.step(1)
    assert o.contact_stats()[0] > 4096
    compare(g, o, sc)


def test_contact_off_leaves_bodies_alone(native):
    sc = scenes.contact_stack(sphere=False)
    sc.contact_degree = 0
    g, o = native.System(sc), refcpu.OracleSystem(sc)
    g.step_forward_in_time(2)
    o.step(2)
    assert g.contact_stats() == (0, 0)
    compare(g, o, sc)


def test_contact_in_a_world_matches_solo(native):
    """Two contact scenes stepped together in one world equal each stepped
    alone, bit for bit (contact runs per scene before its bodies step)."""
    scs = [scenes.contact_stack(), scenes.contact_stack(seed=4, tilt=-0.03)]
    solo = [native.System(s) for s in scs]
    world = native.World([native.System(s) for s in scs])
    for s in solo:
        s.step_forward_in_time(3)
    world.step(3)
    for k, s in enumerate(solo):
        w = world.members[k]
        assert w.contact_stats() == s.contact_stats()
        np.testing.assert_array_equal(bits(w.contact_cells()), bits(s.contact_cells()))
        for b, bd in enumerate(scs[k].boundaries):
            if bd.rigid is not None:
                np.testing.assert_array_equal(bits(w.body_state(b)["com"]),
                                              bits(s.body_state(b)["com"]))
        ws, ss = w.state(), s.state()
        for f in ("x", "v", "p"):
            np.testing.assert_array_equal(bits(ws[f]), bits(ss[f]), err_msg=f)
