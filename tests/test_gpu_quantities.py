"""Sensor queries on the device against the oracle.

sph_quantity_at_positions walks each query's stencil in the reference's
neighbor order and evaluates every term with the oracle's unfused binary64
operations, so the f32 results must be bit-identical to the oracle's, for all
six ScalarQuantity values, with static-sample, volume-map and moving
volume-map boundaries.
"""
import numpy as np
import pytest

import refcpu
from sph_amd import scenes

pytestmark = pytest.mark.gpu

NAMES = ["Speed", "Volume", "Density", "DensityError", "Pressure", "KineticEnergy"]


@pytest.fixture(scope="module")
def native():
    from sph_amd import system
    return system


def same_bits(a, b):
    return a.shape == b.shape and np.array_equal(a.view(np.uint32), b.view(np.uint32))


def queries(x, seed):
    rng = np.random.default_rng(seed)
    lo, hi = x.min(0) - 0.05, x.max(0) + 0.05
    q = [rng.uniform(lo, hi, size=(1500, 3)),            # inside, near walls and outside
         x[rng.choice(len(x), 300, replace=False)],          # at particles (rounded to f32)
         np.array([[50.0, 50.0, 50.0], [-3.0, 0.1, 0.2]])]    # far from everything
    return np.concatenate(q).astype(np.float32)


def splitting_scene():
    sc = scenes.dam_break_sesph(side=10)
    sc.params.solver = "SESPHwSplitting"
    return sc


@pytest.mark.parametrize("case", ["iisph_box", "ost_box", "rigid", "sesph", "splitting"])
def test_quantities_bit_identical(native, case):
    sc = {"iisph_box": lambda: scenes.small_iisph_box(side=12, jitter=2e-3),
          "ost_box": lambda: scenes.small_iisph_box(side=12, solver="IISPHwOST", jitter=2e-3),
          "rigid": scenes.small_rigid_scene,
          "sesph": lambda: scenes.dam_break_sesph(side=10),
          "splitting": splitting_scene}[case]()
    o, g = refcpu.OracleSystem(sc), native.System(sc)
    for step in range(3):
        o.step(2)
        g.step_forward_in_time(2)
        q = queries(g.state()["x"], step)
        print(f"{case} step {2 * step + 2}: {len(q)} queries", flush=True)
        for k, name in enumerate(NAMES):
            go = g.get_quantity_at_positions(name, q)
            oo = o.quantity_at_positions(k, q)
            assert same_bits(go, oo), (name, int(np.sum(go != oo)))
            if name in ("Speed", "Density", "KineticEnergy"):
                assert np.count_nonzero(go) > 100, name


def test_quantities_with_a_moving_sample_body(native):
    """Samples on a moving StaticSampleBoundary body carry the body's
    velocity into the Speed sums: queries right after System::new, where the
    oracle and the device hold the same bits (body forces grouped differently
    part the two at rounding level once the body steps)."""
    sc = scenes.sample_body_scene()
    o, g = refcpu.OracleSystem(sc), native.System(sc)
    q = queries(np.concatenate([g.state()["x"], sc.boundaries[1].rigid.com[None, :]]), 4)
    print(f"moving sample body: {len(q)} queries", flush=True)
    for k, name in enumerate(NAMES):
        go = g.get_quantity_at_positions(name, q)
        oo = o.quantity_at_positions(k, q)
        assert same_bits(go, oo), (name, int(np.sum(go != oo)))


def test_queries_leave_state_and_trajectory_unchanged(native):
    sc = scenes.small_rigid_scene()
    a, b = native.System(sc), native.System(sc)
    a.step_forward_in_time(2)
    b.step_forward_in_time(2)
    q = queries(a.state()["x"], 9)
    for name in NAMES:
        a.get_quantity_at_positions(name, q)
    a.step_forward_in_time(3)
    b.step_forward_in_time(3)
    sa, sb = a.state(), b.state()
    for k in ("x", "v", "a", "V", "p"):
        assert np.array_equal(sa[k], sb[k]), k


def test_empty_and_bad_requests(native):
    g = native.System(scenes.small_iisph_box())
    assert g.get_quantity_at_positions("Density", np.zeros((0, 3))).shape == (0,)
    with pytest.raises(Exception, match="quantity"):
        g.get_quantity_at_positions(6, np.zeros((1, 3)))
