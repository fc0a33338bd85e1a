"""CPU oracle pins (run without a GPU).

Each check pins a part of the oracle (oracle/refcpu.c) with a known answer
from the reference's own tests (cited file:line, values in
tests/golden/reference_known_answers.json), a closed form, an invariant or a
brute-force computation, chosen so that a dropped term, a wrong sign or index
or a transposed operand fails.
"""
import json
import math
import os

import numpy as np
import pytest

import refcpu
from sph_amd import scenes

HERE = os.path.dirname(os.path.abspath(__file__))
with open(os.path.join(HERE, "golden", "reference_known_answers.json")) as f:
    KA = json.load(f)

L = refcpu.lib()


# ------------------------------------------------------------------ kernel
def test_kernel_known_answers():
    # tests/sph/kernel/cubic_spline.rs:16-30, 90-96
    for q, w in KA["kernel_w"]:
        assert abs(L.ref_w(q) - w) < 1e-12
    for q in (1.0, 1.5, 10.0):
        assert L.ref_w(q) == 0.0
        assert L.ref_dqw(q) == 0.0 or q < 1.0


def test_kernel_derivative_matches_finite_difference():
    # tests/sph/kernel/cubic_spline.rs:70-82
    h = 1e-6
    for q in (0.05, 0.2, 0.35, 0.55, 0.7, 0.9):
        num = (L.ref_w(q + h) - L.ref_w(q - h)) / (2 * h)
        assert abs(L.ref_dqw(q) - num) < 1e-4


def test_kernel_normalization_and_gradient():
    # integral of W over its support is 1 (tests/sph/kernel/cubic_spline.rs:136-160)
    h = 0.7
    xs, ws = np.polynomial.legendre.leggauss(64)
    r = 0.5 * h * (xs + 1)
    wr = 0.5 * h * ws
    vals = np.array([refcpu.kernel_function([ri, 0.0, 0.0], h) for ri in r])
    assert abs(np.sum(4 * math.pi * r * r * vals * wr) - 1.0) < 1e-6
    # gradient is exactly zero at r = 0 and beyond the support (:190-205)
    assert np.array_equal(refcpu.kernel_gradient([0, 0, 0], 1.0), np.zeros(3))
    assert np.array_equal(refcpu.kernel_gradient([1.5, 0, 0], 1.0), np.zeros(3))
    # gradient = finite difference of W, antisymmetric in r
    rv = np.array([0.21, -0.13, 0.05])
    g = refcpu.kernel_gradient(rv, h)
    e = 1e-7
    fd = [(refcpu.kernel_function(rv + e * np.eye(3)[k], h) -
           refcpu.kernel_function(rv - e * np.eye(3)[k], h)) / (2 * e) for k in range(3)]
    assert np.allclose(g, fd, rtol=1e-5, atol=1e-6)
    assert np.array_equal(refcpu.kernel_gradient(-rv, h), -g)


# ------------------------------------------------------------------ hashing
def test_spatial_hash_known_answers():
    # neighbor_search/spatial_hashing.rs:173-247
    for p, cs, cell in KA["get_cell"]:
        assert refcpu.get_cell(p, cs) == tuple(cell)
    for k, u in KA["map_int_to_uint"]:
        assert L.ref_map_int_to_uint(k) == u


def test_hash_is_collision_free_on_the_test_range():
    # spatial_hashing.rs:251-275: no collisions over [-20, 20]^3
    seen = set()
    for k in range(-20, 21):
        for l in range(-20, 21):
            for m in range(-20, 21):
                seen.add(L.ref_hash(k, l, m))
    assert len(seen) == 41 ** 3


def brute_force(pos, smp, rng):
    out = []
    for p in pos:
        d = np.sqrt(np.sum((p - smp) ** 2, axis=1))
        out.append(sorted(np.nonzero(d < rng)[0].tolist()))
    return out


@pytest.mark.parametrize("cs,rng", [(0.1, 0.1), (0.05, 0.1), (0.1, 0.07)])
def test_neighbor_sets_equal_brute_force(cs, rng):
    r = np.random.default_rng(3)
    pos = r.uniform(-0.4, 0.4, size=(400, 3))
    smp = r.uniform(-0.4, 0.4, size=(500, 3))
    off, idx = refcpu.find_samples(cs, rng, pos, smp)
    bf = brute_force(pos, smp, rng)
    for i in range(len(pos)):
        assert sorted(idx[off[i]:off[i + 1]].tolist()) == bf[i]


def test_neighbor_order_and_strict_range():
    # self included, strict < (spatial_hashing.rs:353-363); symmetric sets
    pos = np.array([[0.0, 0, 0], [2.0, 0, 0], [1.0, 0, 0]])
    off, idx = refcpu.find_samples(1.0, 2.0, pos, pos)
    assert idx[off[0]:off[1]].tolist() == [0, 2]
    assert 0 not in idx[off[1]:off[2]].tolist()
    sc = scenes.small_iisph_box(side=6)
    off, idx = refcpu.find_samples(0.04, 0.04, sc.x, sc.x)
    sets = [set(idx[off[i]:off[i + 1]].tolist()) for i in range(sc.n)]
    for i, s in enumerate(sets):
        assert i in s
        for j in s:
            assert i in sets[j]


# ------------------------------------------------------------------ serendipity
def test_reference_nodes_and_offsets():
    nodes = np.zeros(96)
    L.ref_reference_nodes(refcpu.dptr(nodes))
    nodes = nodes.reshape(32, 3)
    assert len({tuple(n) for n in nodes}) == 32
    offs = np.array([[L.ref_node_offset(c) for c in n] for n in nodes])
    assert np.array_equal(offs, scenes.NODE_OFFSETS)   # discretization.rs:567-573


def test_shape_functions_kronecker_and_partition_of_unity():
    # discretization.rs:640-681
    nodes = np.zeros(96)
    L.ref_reference_nodes(refcpu.dptr(nodes))
    nodes = nodes.reshape(32, 3)
    out = np.zeros(32)
    for k, n in enumerate(nodes):
        L.ref_shape_functions(n[0], n[1], n[2], refcpu.dptr(out))
        exp = np.zeros(32)
        exp[k] = 1.0
        assert np.allclose(out, exp, atol=1e-12)
    r = np.random.default_rng(0)
    for _ in range(20):
        xi, eta, zeta = r.uniform(-1, 1, 3)
        L.ref_shape_functions(xi, eta, zeta, refcpu.dptr(out))
        assert abs(out.sum() - 1.0) < 1e-12


def _lattice(fn, lo=(-0.3, -0.3, -0.3), hi=(0.3, 0.3, 0.3), dx=0.1):
    lo, hi = np.array(lo), np.array(hi)
    n = np.round((hi - lo) / dx).astype(np.int64)
    pts = scenes._node_positions(lo, n, dx)
    vals = fn(pts)
    return scenes.Lattice(lo, hi, dx, n, np.ascontiguousarray(vals),
                          np.ones(vals.shape, dtype=np.uint8))


def test_serendipity_reproduces_linear_fields_and_gradients():
    # the cubic serendipity element reproduces affine functions exactly
    a = np.array([0.3, -1.2, 2.0])
    lat = _lattice(lambda p: p @ a + 0.7)
    r = np.random.default_rng(1)
    for p in r.uniform(-0.25, 0.25, size=(30, 3)):
        st, v = refcpu.serendipity_function(lat, p)
        assert st == 0 and abs(v - (p @ a + 0.7)) < 1e-12
        st, g = refcpu.serendipity_gradient(lat, p)
        assert st == 0 and np.allclose(g, a, atol=1e-9)


def test_serendipity_bounds_and_pruning():
    lat = _lattice(lambda p: p[..., 0])
    # outside the domain -> OutOfBounds (discretization.rs:780-800)
    assert refcpu.serendipity_function(lat, [0.31, 0, 0])[0] == 1
    # exactly on the upper boundary -> clamped into the last cell (:762-778)
    st, v = refcpu.serendipity_function(lat, [0.3, 0.0, 0.0])
    assert st == 0 and abs(v - 0.3) < 1e-12
    # a missing node of the cell -> PrunedCell (:836-865)
    lat.present[0, 0, 0] = 0
    assert refcpu.serendipity_function(lat, [-0.29, -0.29, -0.29])[0] == 2
    # gradient near the pruned cell propagates the error without fallback (:886-920)
    assert refcpu.serendipity_gradient(lat, [-0.19, -0.29, -0.29])[0] != 0


# ------------------------------------------------------------------ rigid body
def _rigid(mass=2.0, w=(0, 0, 0)):
    rb = refcpu.C.create_string_buffer(1024)
    I = np.eye(3)
    L.ref_rigid_init.argtypes = [refcpu.C.c_void_p] + [refcpu._dp] * 0 + [
        refcpu.C.c_double, refcpu._dp, refcpu._dp, refcpu._dp, refcpu._dp, refcpu._dp,
        refcpu._dp]
    arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in
            (I, I, np.zeros(3), np.array([1.0, 0, 0, 0]), np.zeros(3), np.array(w, float))]
    L.ref_rigid_init(rb, mass, *[refcpu.dptr(a) for a in arrs])
    return rb, arrs


def _state(rb):
    # ref_rigid layout: mass, Iinv_body[9], com[3], q[4], v[3], L[3], F[3], T[3], ...
    a = np.frombuffer(rb.raw[:8 * 41], dtype=np.float64)
    return {"com": a[10:13], "q": a[13:17], "v": a[17:20], "F": a[23:26], "T": a[26:29],
            "w": a[38:41]}


def test_rigid_body_euler_cromer_and_rotation():
    # tests/sph/boundary_handling/rigid_body_motion.rs:393-470
    L.ref_rigid_add_force.argtypes = [refcpu.C.c_void_p, refcpu._dp, refcpu._dp]
    L.ref_rigid_step.argtypes = [refcpu.C.c_void_p, refcpu.C.c_double]
    rb, _ = _rigid(2.0)
    f = np.array([4.0, 0, 0])
    L.ref_rigid_add_force(rb, refcpu.dptr(f), refcpu.dptr(np.zeros(3)))
    L.ref_rigid_step(rb, 0.1)
    s = _state(rb)
    assert np.allclose(s["v"], KA["rigid_linear_step"]["velocity"], atol=1e-12)
    assert np.allclose(s["com"], KA["rigid_linear_step"]["position"], atol=1e-12)
    assert np.allclose(s["F"], 0) and np.allclose(s["T"], 0)
    # torque-free spin about z conserves omega; orientation turns by omega dt
    rb, _ = _rigid(1.0, (0, 0, 2.0))
    L.ref_rigid_step(rb, 0.001)
    s = _state(rb)
    ang = 2 * math.atan2(np.linalg.norm(s["q"][1:]), s["q"][0])
    assert abs(ang - 0.002) < 1e-12 and abs(abs(s["q"][3]) - math.sin(0.001)) < 1e-12


# ------------------------------------------------------------------ system pins
def _single_particle_scene(q=0.25, h=1.0, dx=0.9, omega=0.5, dt=0.05):
    """One fluid particle and one static boundary sample at distance q h; radial
    gravity centred on the particle (zero acceleration)."""
    x = np.zeros((1, 3))
    p = scenes.Params(kernel_support_radius=h, rest_density_grid_spacing=dx, time_increment=dt,
                      fluid_viscosity=0.0, boundary_viscosity=0.0, solver="IISPH",
                      integrator="EulerCromer", gravity_mode=1, relaxation_factor=omega,
                      boundary_pressure_acceleration_weighting=0.7,
                      target_density_error=1e-3)
    b = scenes.StaticSamples(np.array([[q * h, 0.0, 0.0]]))
    return scenes.Scene("single", p, x, np.zeros((1, 3)), np.array([0.5]), np.zeros(1, np.uint32), [b])


def test_iisph_single_particle_closed_form():
    """For one particle with one boundary neighbor the Jacobi map is
    p <- p + w (s/a - p), so p_k = (s/a)(1 - (1-w)^(k+1)) in closed form
    (iisph.rs:344-376, 401-560; tests/sph/pressure_solver/iisph.rs:431-490)."""
    q, h, dx, om, dt, gam = 0.25, 1.0, 0.9, 0.5, 0.05, 0.7
    sc = _single_particle_scene(q, h, dx, om, dt)
    o = refcpu.OracleSystem(sc)
    st, m = o.state(), o.measure()
    W0 = 16 / math.pi / h ** 3 * 0.5
    Vb = 1.0 / W0                                    # pseudo volume of a lone sample
    wq = (1 - q) ** 3 - 4 * (0.5 - q) ** 3
    V0 = dx ** 3
    V = V0 / (V0 * W0 + Vb * 16 / math.pi / h ** 3 * wq)
    assert abs(st["V"][0] - V) <= 1e-12 * V
    dqw = -3 * (1 - q) ** 2 + 12 * (0.5 - q) ** 2
    grad = -16 / math.pi / h ** 4 * dqw              # d/dx of W at r = -q h e_x
    Sb = Vb * grad
    s = 1 - V0 / V
    a = dt ** 2 * (-(V / 0.5) * (2 * gam * Sb)) * Sb
    assert s < 0 and a < 0
    # iteration t sees p_(t-1) = (s/a)(1 - (1-w)^t), so its error is 100 |s| (1-w)^t
    t = 0
    while True:
        t += 1
        err = 100 * abs(s) * (1 - om) ** t
        if t > 2 and err < 1e-3:
            break
    assert m["solver_iterations"] == t
    p_exp = (s / a) * (1 - (1 - om) ** (t + 1))
    assert abs(st["p"][0] - p_exp) <= 1e-9 * abs(p_exp)


def test_sesph_pressure_is_linear_in_stiffness_and_conserves_momentum():
    sc = scenes.small_iisph_box(side=8, jitter=3e-3, solver="SESPH")
    sc.boundaries = []
    sc.params.fluid_viscosity = 0.0
    sc.params.stiffness = 1000.0
    a1 = refcpu.OracleSystem(sc).state()
    sc.params.stiffness = 2000.0
    a2 = refcpu.OracleSystem(sc).state()
    sc.params.stiffness = 0.0
    a0 = refcpu.OracleSystem(sc).state()
    assert np.allclose(a2["p"], 2 * a1["p"], rtol=1e-14, atol=0)
    f1 = a1["a"] - a0["a"]
    f2 = a2["a"] - a0["a"]
    assert np.allclose(f2, 2 * f1, rtol=1e-9, atol=1e-12)
    assert np.max(np.abs(f1)) > 1e-3
    # pairwise forces are antisymmetric: total pressure force vanishes
    tot = (a1["m"][:, None] * f1).sum(axis=0)
    assert np.max(np.abs(tot)) <= 1e-9 * np.sum(a1["m"][:, None] * np.abs(f1))


def test_volume_of_rest_lattice_is_near_rest_volume():
    # quantities.rs:610-633: interior lattice particle within 15 % of V0
    dx = 0.05
    g = np.arange(-4, 5) * dx
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    x = np.stack([X.ravel(), Y.ravel(), Z.ravel()], -1)
    p = scenes.Params(kernel_support_radius=4 * dx, rest_density_grid_spacing=dx,
                      solver="SESPH")
    sc = scenes.Scene("lattice", p, x, np.zeros_like(x), np.full(len(x), 1000 * dx ** 3),
                      np.zeros(len(x), np.uint32), [])
    st = refcpu.OracleSystem(sc).state()
    c = int(np.argmin(np.sum(x ** 2, axis=1)))
    assert abs(st["V"][c] - dx ** 3) / dx ** 3 < 0.15


def test_compaction_swap_order():
    # Fluid::disable swaps with the last active particle (fluid.rs:159-188):
    # z = [-1, 2, -3, 4, 5], threshold 0 -> surviving originals [4, 1, 3]
    x = np.array([[0.0, 0, z] for z in KA["compaction"]["z"]])
    p = scenes.Params(kernel_support_radius=0.1, rest_density_grid_spacing=0.05,
                      solver="SESPH", disable_particles_below=0.0)
    sc = scenes.Scene("compact", p, x, np.zeros_like(x), np.ones(len(x)),
                      np.arange(len(x), dtype=np.uint32), [])
    st = refcpu.OracleSystem(sc).state()
    assert st["fluid_id"].tolist() == KA["compaction"]["survivors"]


def test_euler_cromer_position_uses_new_velocity():
    # tests/sph/integration_schemes/euler_cromer.rs:91-132 on an isolated particle
    x = np.array([[10.0, 0, 0]])
    v = np.array([[1.0, 0.5, -0.25]])
    dt = 0.01
    p = scenes.Params(kernel_support_radius=0.1, rest_density_grid_spacing=0.05,
                      time_increment=dt, solver="SESPH")
    sc = scenes.Scene("free", p, x, v, np.ones(1), np.zeros(1, np.uint32), [])
    o = refcpu.OracleSystem(sc)
    o.step(1)
    st = o.state()
    g = np.array([0, 0, -9.81])
    v1 = v[0] + dt * g
    assert np.allclose(st["v"][0], v1, rtol=0, atol=1e-15)
    assert np.allclose(st["x"][0], x[0] + dt * v1, rtol=0, atol=1e-15)


def test_measurement_reductions():
    sc = scenes.small_iisph_box(side=6)
    o = refcpu.OracleSystem(sc)
    st, m = o.state(), o.measure()
    V0 = sc.params.rest_density_grid_spacing ** 3
    dens = np.where(st["V"] < V0, st["m"] / st["V"], st["m"] / V0).mean()
    err = 100 * (np.where(st["V"] < V0, V0 / st["V"], 1.0).mean() - 1)
    assert abs(m["density"] - dens) <= 1e-12 * dens
    assert abs(m["density_error"] - err) <= 1e-9


def test_cfl_time_step():
    # set_cfl_time_step (sph/mod.rs:825-829): v = 0 -> dt_max; else min(dt_max, lambda dx / v)
    sc = scenes.small_iisph_box(side=4)
    o = refcpu.OracleSystem(sc)
    assert o.measure()["time_step_size"] == sc.params.max_time_increment
    sc.v[:] = 0.0
    sc.v[0] = [0.0, 30.0, 0.0]
    o = refcpu.OracleSystem(sc)
    exp = 0.4 * sc.params.rest_density_grid_spacing / 30.0
    assert abs(o.measure()["time_step_size"] - min(1e-3, exp)) < 1e-18
