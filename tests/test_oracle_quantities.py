"""CPU pins of the oracle's sensor queries (ref_system_quantity_at_positions,
following SPHSystem::get_quantity_at_positions, sph/mod.rs:241-341, and
quantities.rs:14-330).

Checks: brute-force O(N*M) SPH sums in numpy with an independent kernel
(scenes.kernel_w) for the fluid-only quantities; the volume at particle
positions against the volume the step itself computed (same neighbors, same
boundary samples); the reference's density-error call, which passes the fluid
masses as sample volumes, on a fluid whose masses are below the rest volume.
"""
import math

import numpy as np
import pytest

import refcpu
from sph_amd import scenes


def brute(x_eval, xs, h):
    """W(|x_eval - xs|, h) for every pair within the support (strict)."""
    d = np.linalg.norm(x_eval[:, None, :] - xs[None, :, :], axis=-1)
    w = 16.0 / math.pi / h ** 3 * scenes.kernel_w(d / h)
    return np.where(d < h, w, 0.0)


@pytest.fixture(scope="module")
def box():
    sc = scenes.small_iisph_box()
    o = refcpu.OracleSystem(sc)
    o.step(3)
    st = o.state()
    rng = np.random.default_rng(5)
    lo, hi = st["x"].min(0) - 0.03, st["x"].max(0) + 0.03
    q = rng.uniform(lo, hi, size=(400, 3)).astype(np.float32)
    return sc, o, st, q


def test_fluid_sums_equal_brute_force(box):
    sc, o, st, q = box
    h = sc.params.kernel_support_radius
    W = brute(q.astype(np.float64), st["x"], h)
    v2 = np.sum(st["v"] ** 2, axis=1)
    want = {
        "Density": W @ st["m"],
        "Pressure": W @ (st["p"] * st["V"]),
        "KineticEnergy": W @ (0.5 * st["m"] * v2 * st["V"]),
    }
    for name, ref in want.items():
        got = o.quantity_at_positions(["Speed", "Volume", "Density", "DensityError", "Pressure",
                                       "KineticEnergy"].index(name), q)
        np.testing.assert_allclose(got, ref.astype(np.float32), rtol=2e-6,
                                   atol=1e-6 * np.abs(ref).max(), err_msg=name)
        assert np.count_nonzero(got) > 100
    # a static volume-map container: boundary samples carry zero velocity
    sp = o.quantity_at_positions(0, q)
    ref = np.linalg.norm(W @ (st["v"] * st["V"][:, None]), axis=1)
    np.testing.assert_allclose(sp, ref.astype(np.float32), rtol=2e-6, atol=1e-6 * ref.max())


def test_volume_at_particle_positions_equals_step_volume():
    # f32-representable positions: queried at the particles themselves, the
    # volume sums see the same neighbors and boundary samples as get_volume
    for sc in (scenes.dam_break_sesph(side=8), ):
        o = refcpu.OracleSystem(sc)            # System::new computes V
        st = o.state()
        assert np.array_equal(st["x"].astype(np.float32).astype(np.float64), st["x"])
        got = o.quantity_at_positions(1, st["x"].astype(np.float32))
        assert np.array_equal(got, st["V"].astype(np.float32))
    sc = scenes.small_iisph_box(jitter=0.0)
    sc.x = sc.x.astype(np.float32).astype(np.float64)
    o = refcpu.OracleSystem(sc)
    st = o.state()
    got = o.quantity_at_positions(1, st["x"].astype(np.float32))
    assert np.array_equal(got, st["V"].astype(np.float32))


def test_density_error_uses_masses_as_volumes():
    sc = scenes.small_iisph_box()
    V0 = sc.params.rest_density_grid_spacing ** 3
    # masses below the rest volume (rest density < 1) make the terms non-zero
    sc.m = sc.m * (0.5 * V0 / sc.m.max())
    o = refcpu.OracleSystem(sc)
    st = o.state()
    q = st["x"][:50].astype(np.float32)
    W = brute(q.astype(np.float64), st["x"], sc.params.kernel_support_radius)
    m = st["m"]
    ref = 100.0 * (W @ np.where(m < V0, (V0 / m - 1.0) * m, 0.0))
    got = o.quantity_at_positions(3, q)
    np.testing.assert_allclose(got, ref.astype(np.float32), rtol=2e-6)
    assert np.all(got > 0)
    # with real masses (rho0 >= 800, so m > V0) every term is skipped
    o2 = refcpu.OracleSystem(scenes.small_iisph_box())
    assert not np.any(o2.quantity_at_positions(3, q))


def test_queries_leave_the_state_unchanged(box):
    sc, _, _, q = box
    a, b = refcpu.OracleSystem(sc), refcpu.OracleSystem(sc)
    a.step(2)
    for k in range(6):
        a.quantity_at_positions(k, q)
    a.step(2)
    b.step(4)
    sa, sb = a.state(), b.state()
    for k in ("x", "v", "V", "p"):
        assert np.array_equal(sa[k], sb[k]), k
