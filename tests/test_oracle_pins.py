"""Pins of the oracle's per-function arithmetic (CPU only).

Each test restates one of the reference's own per-function tests (given
volumes, pressures, velocities, hand-derived expectations), or checks a
property the mathematics fixes (consistency of the SPH operators on a
lattice, momentum balance, kinematic identities, textbook rigid-body torque),
against the oracle's single reference functions (the refcpu.h test hooks).
They are chosen so that a dropped factor, a flipped sign or a swapped operand
in the oracle fails at least one of them.

The kernel gradient used for the expectations is the cubic B-spline's closed
form written out here (W = 16/pi h^-3 w(q), grad W = 16/pi h^-4 r/|r| w'(q);
kernel/mod.rs:22-46, cubic_spline.rs:8-17), itself pinned in test_oracle.py.
"""
import math

import numpy as np
import pytest

import refcpu
from sph_amd import scenes

ALPHA = 16.0 / math.pi


def grad_w(r, h):
    r = np.asarray(r, dtype=np.float64)
    d = math.sqrt(float(r @ r))
    if d == 0.0:
        return np.zeros(3)
    q = d / h
    dw = -3.0 * max(1.0 - q, 0.0) ** 2 + 12.0 * max(0.5 - q, 0.0) ** 2
    return ALPHA / h ** 4 * r / d * dw


def kernel(r, h):
    q = math.sqrt(float(np.asarray(r) @ np.asarray(r))) / h
    return ALPHA / h ** 3 * (max(1.0 - q, 0.0) ** 3 - 4.0 * max(0.5 - q, 0.0) ** 3)


def make(x, v=None, m=None, h=1.0, dx=0.3, boundaries=(), **kw):
    """An oracle system on the given particles (SESPH for the set-up update;
    the tests then call single functions on overwritten fields)."""
    x = np.asarray(x, dtype=np.float64).reshape(-1, 3)
    n = len(x)
    v = np.zeros((n, 3)) if v is None else np.asarray(v, dtype=np.float64)
    m = np.full(n, 0.5) if m is None else np.asarray(m, dtype=np.float64)
    kw.setdefault("solver", "SESPH")
    P = scenes.Params(kernel_support_radius=h, rest_density_grid_spacing=dx, **kw)
    sc = scenes.Scene("pin", P, x, v, m, np.zeros(n, dtype=np.uint32), list(boundaries),
                      contact_degree=0)
    return refcpu.OracleSystem(sc)


def neighbors_of(o, i):
    off, idx = o.fluid_neighbors()
    return [int(j) for j in idx[off[i]:off[i + 1]]]


def lattice(ratio, dx=0.02):
    """(2 ratio + 3)^3 lattice points at spacing dx around the origin."""
    k = np.arange(-ratio - 1, ratio + 2, dtype=np.float64)
    X = np.stack(np.meshgrid(k, k, k, indexing="ij"), -1).reshape(-1, 3) * dx
    centre = int(np.argmin(np.sum(X * X, axis=1)))
    return X, centre


def dynamic_body(com):
    return scenes.Rigid(mass=2.0, inertia_body=np.diag([1.0, 2.0, 3.0]),
                        com=np.asarray(com, dtype=np.float64))


# ------------------------------------------------------------------ viscosity
POS3 = np.array([[0.0, 0.0, 0.0], [0.3, 0.0, 0.0], [0.0, 0.3, 0.0]])


def test_viscosity_matches_the_reference_manual_formula():
    """non_pressure_accelerations.rs:685-727: three particles with given
    volumes; a_i = sum_j nu 2 (d+2) V_j ((v_i-v_j).(x_i-x_j)) /
    (|x_i-x_j|^2 + 0.01 dx^2) grad W(x_i - x_j)."""
    vel = np.array([[1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0]])
    vols = np.array([0.02, 0.025, 0.03])
    nu, h, dx = 5.0, 1.0, 0.3
    o = make(POS3, vel, h=h, dx=dx, fluid_viscosity=nu, boundary_viscosity=nu)
    o.set_field("V", vols)
    o.call("reset_acceleration")
    o.call("viscosity")
    a = o.field("a")
    for i in range(3):
        exp = np.zeros(3)
        for j in neighbors_of(o, i):
            r = POS3[i] - POS3[j]
            exp += (nu * 2.0 * (3.0 + 2.0) * vols[j] * float((vel[i] - vel[j]) @ r)
                    / (float(r @ r) + 0.01 * dx * dx) * grad_w(r, h))
        assert np.linalg.norm(a[i] - exp) <= 1e-12 * max(1.0, np.linalg.norm(exp)), i
        assert np.linalg.norm(exp) > 0.1


def test_viscosity_is_consistent_with_nu_laplacian_on_a_lattice():
    """The factor 2(d+2) = 10 makes the SPH viscosity term consistent with
    nu * laplacian(v) for a divergence-free field: with v = (y^2 + z^2, 0, 0)
    an interior particle of a lattice (h = 4 dx, V = dx^3) gets a_x = 4 nu
    within the discretization error (~1.3 % at this resolution)."""
    X, c = lattice(4)
    dx, nu = 0.02, 0.01
    v = np.zeros_like(X)
    v[:, 0] = X[:, 1] ** 2 + X[:, 2] ** 2
    o = make(X, v, m=np.full(len(X), 1000 * dx ** 3), h=4 * dx, dx=dx, fluid_viscosity=nu)
    o.set_field("V", np.full(len(X), dx ** 3))
    o.call("reset_acceleration")
    o.call("viscosity")
    a = o.field("a")[c]
    assert abs(a[0] / (4 * nu) - 1.0) < 0.03, a
    assert abs(a[1]) < 1e-12 and abs(a[2]) < 1e-12


def test_identical_velocities_give_no_viscosity():
    """non_pressure_accelerations.rs:729-764."""
    vel = np.tile([2.0, -1.0, 0.5], (2, 1))
    o = make(POS3[:2], vel, fluid_viscosity=5.0)
    o.set_field("V", [0.02, 0.02])
    o.call("reset_acceleration")
    o.call("viscosity")
    assert np.all(np.abs(o.field("a")) < 1e-12)


# -------------------------------------------------------------------- gravity
def test_radial_and_uniform_gravity():
    """non_pressure_accelerations.rs:64-103: uniform a = (0, 0, -9.81); radial
    a = 9.81 (c - x) / sqrt(|c - x|^2 + 1e-4) (unit magnitude direction with
    softening); a dynamic body gets M g at its centre of mass (no torque)."""
    x = np.array([[3.0, 0.0, 4.0], [0.0, 0.0, 0.0], [0.0, 0.01, 0.0]])
    o = make(x, gravity_mode=1, gravity_center=(0.0, 0.0, 0.0))
    o.call("reset_acceleration")
    o.call("gravity")
    a = o.field("a")
    assert np.allclose(a[0], 9.81 * np.array([-3.0, 0.0, -4.0]) / math.sqrt(25.0 + 1e-4),
                       rtol=1e-15, atol=0)
    assert np.all(a[1] == 0.0)
    assert np.allclose(a[2], [0.0, -9.81 * 0.01 / math.sqrt(2e-4), 0.0], rtol=1e-15, atol=0)
    b = scenes.StaticSamples(positions=np.array([[0.0, 0.0, 0.0]]),
                             rigid=dynamic_body([0.0, 0.0, 50.0]))
    o = make(x[:1], boundaries=[b])
    o.call("neighbors")
    o.call("reset_acceleration")
    o.call("gravity")
    assert np.array_equal(o.field("a")[0], [0.0, 0.0, -9.81])
    assert np.array_equal(o.body_forces(0), [0.0, 0.0, -9.81 * 2.0, 0.0, 0.0, 0.0])


# ------------------------------------------------------ pressure acceleration
def test_pressure_acceleration_matches_the_reference_manual_formula():
    """pressure_solver/mod.rs:535-588 with unequal masses: accumulates onto the
    existing acceleration; a_i -= V_i/m_i V_j (p_i + p_j) grad W_ij."""
    p = np.array([100.0, 150.0, 80.0])
    vols = np.array([0.02, 0.025, 0.03])
    m = np.array([0.5, 0.7, 1.3])
    o = make(POS3, m=m)
    pre = np.tile([0.0, 0.0, -9.81], (3, 1))
    o.set_field("V", vols)
    o.set_field("p", p)
    o.set_field("a", pre)
    o.call("pressure_acceleration")
    a = o.field("a")
    for i in range(3):
        exp = pre[i].copy()
        for j in neighbors_of(o, i):
            exp -= vols[i] / m[i] * vols[j] * (p[i] + p[j]) * grad_w(POS3[i] - POS3[j], 1.0)
        assert np.linalg.norm(a[i] - exp) <= 1e-12 * np.linalg.norm(exp), i
    o.call("pressure_acceleration_overwrite")  # overwrite discards the existing acceleration
    assert np.allclose(o.field("a"), a - pre, rtol=0, atol=1e-12)


def test_pressure_force_conserves_momentum_with_unequal_masses():
    """Pairwise antisymmetry: sum_i m_i a_i = 0 for the fluid-fluid pressure
    acceleration of any particle set, for any masses and pressures."""
    rng = np.random.default_rng(3)
    x = rng.uniform(0.0, 0.12, size=(200, 3))
    m = rng.uniform(0.5, 2.0, size=200)
    o = make(x, m=m, h=0.04, dx=0.02)
    o.set_field("V", rng.uniform(6e-6, 1e-5, size=200))
    o.set_field("p", rng.uniform(0.0, 500.0, size=200))
    o.call("pressure_acceleration_overwrite")
    a = o.field("a")
    f = m[:, None] * a
    assert np.linalg.norm(f.sum(axis=0)) <= 1e-12 * np.abs(f).sum()
    assert np.abs(f).sum() > 1.0


def test_pressure_acceleration_is_consistent_with_the_pressure_gradient():
    """For a linear pressure field p = p0 + G z on a lattice (h = 4 dx,
    V = dx^3, m = rho0 dx^3) the symmetric SPH form gives a = -grad(p)/rho0
    at an interior particle (to 1e-4 at this resolution)."""
    X, c = lattice(4)
    dx, rho0, G = 0.02, 1000.0, 3000.0
    o = make(X, m=np.full(len(X), rho0 * dx ** 3), h=4 * dx, dx=dx)
    o.set_field("V", np.full(len(X), dx ** 3))
    o.set_field("p", 1.0e4 + G * X[:, 2])
    o.call("pressure_acceleration_overwrite")
    a = o.field("a")[c]
    assert abs(a[2] / (-G / rho0) - 1.0) < 1e-3, a
    assert abs(a[0]) < 1e-9 and abs(a[1]) < 1e-9


def one_static_sample(pos, rigid=None):
    return scenes.StaticSamples(positions=np.asarray([pos], dtype=np.float64), rigid=rigid)


def test_boundary_pressure_acceleration_matches_the_reference_formula():
    """pressure_solver/mod.rs:668-723: one static sample at (0.2, 0, 0) with
    volume 0.01, weighting 0.5: a = -2 w V_i V_b p_i grad W / m_i."""
    w = 0.5
    o = make([[0.0, 0.0, 0.0]], m=[0.5], boundaries=[one_static_sample([0.2, 0.0, 0.0])],
             boundary_pressure_acceleration_weighting=w)
    o.sample_volumes(0, [0.01])
    o.set_field("V", [0.02])
    o.set_field("p", [200.0])
    o.call("pressure_acceleration_overwrite")
    force = 2.0 * w * 0.02 * 0.01 * 200.0 * grad_w(np.array([-0.2, 0.0, 0.0]), 1.0)
    exp = -force / 0.5
    a = o.field("a")[0]
    assert np.linalg.norm(a - exp) <= 1e-13 * np.linalg.norm(exp)
    assert np.all(o.body_forces(0) == 0.0)  # a static boundary takes no reaction


def test_pressure_reaction_on_a_dynamic_boundary_and_its_torque():
    """pressure_solver/mod.rs:724-775: the force recorded on a dynamic boundary
    is minus the force felt by the fluid, at the sample position; the body's
    torque about its centre of mass is (x_b - c) x F (rigid_body_motion.rs:
    143-146)."""
    com = np.array([5.0, 0.0, 0.0])
    # body-frame sample at (0.2,0,0) - com, so that the world position is (0.2, 0, 0)
    b = scenes.StaticSamples(positions=np.array([[0.2 - 5.0, 0.1, 0.0]]),
                             rigid=dynamic_body(com))
    o = make([[0.0, 0.0, 0.0]], m=[0.5], boundaries=[b])
    o.sample_volumes(0, [0.01])
    o.set_field("V", [0.02])
    o.set_field("p", [150.0])
    o.call("neighbors")
    o.call("pressure_acceleration_overwrite")
    f_fluid = 0.5 * o.field("a")[0]
    F = o.body_forces(0)
    assert np.linalg.norm(F[:3] + f_fluid) <= 1e-13 * np.linalg.norm(f_fluid)
    xb = np.array([0.2, 0.1, 0.0])
    assert np.allclose(F[3:], np.cross(xb - com, F[:3]), rtol=1e-13, atol=0)
    assert abs(F[5]) > 1e-3


def test_viscosity_reaction_on_a_dynamic_boundary():
    """non_pressure_accelerations.rs:877-925: a dynamic boundary sample takes
    -m_i times the fluid's boundary viscosity acceleration, at the sample."""
    com = np.array([5.0, 0.0, 0.0])
    b = scenes.StaticSamples(positions=np.array([[0.25 - 5.0, 0.0, 0.0]]),
                             rigid=dynamic_body(com))
    o = make([[0.0, 0.0, 0.0]], v=[[1.0, 0.5, 0.0]], m=[0.5], boundaries=[b],
             boundary_viscosity=2.0)
    o.sample_volumes(0, [0.01])
    o.call("neighbors")
    o.call("reset_acceleration")
    o.call("viscosity")
    a = o.field("a")[0]
    r = np.array([-0.25, 0.0, 0.0])
    exp = (2.0 * 10.0 * 0.01 * float(np.array([1.0, 0.5, 0.0]) @ r)
           / (float(r @ r) + 0.01 * 0.09) * grad_w(r, 1.0))
    assert np.linalg.norm(a - exp) <= 1e-13 * np.linalg.norm(exp)
    F = o.body_forces(0)
    assert np.allclose(F[:3], -0.5 * exp, rtol=1e-13, atol=0)


# --------------------------------------------------------------------- SESPH
def test_sesph_state_equation():
    """sesph.rs:310-441: p = k max(V0/V - 1, 0) per particle; zero at or above
    the rest volume."""
    k, dx = 800.0, 0.3
    V0 = (dx * dx) * dx  # rest_volume = dx.powi(3) (sph/mod.rs:816)
    far = np.array([[0.0, 0.0, 0.0], [1000.0, 0.0, 0.0], [0.0, 1000.0, 0.0], [0, 0, 1000.0]])
    o = make(far, dx=dx, stiffness=k)
    vols = np.array([0.8 * V0, 0.5 * V0, V0, 1.5 * V0])
    o.set_field("V", vols)
    o.call("reset_acceleration")
    o.call("sesph")
    p = o.field("p")
    assert abs(p[0] - k * (1 / 0.8 - 1)) <= 1e-12 * k
    assert abs(p[1] - k * (1 / 0.5 - 1)) <= 1e-12 * k
    assert p[2] == 0.0 and p[3] == 0.0


def test_sesph_pressure_of_a_compressed_lattice():
    """A lattice at spacing s = 0.9 dx: get_volume's V0 / V is the kernel sum
    V0 sum_j W_ij = (dx/s)^3 at an interior particle (partition of unity of the
    kernel on a lattice, to 1e-4 for h = 4 s), so p = k ((dx/s)^3 - 1)."""
    s, dx, k = 0.018, 0.02, 1000.0
    X, c = lattice(4, dx=s)
    o = make(X, m=np.full(len(X), 1000 * dx ** 3), h=4 * s, dx=dx, stiffness=k)
    o.call("volume")
    o.call("reset_acceleration")
    o.call("sesph")
    p = o.field("p")[c]
    exp = k * ((dx / s) ** 3 - 1.0)
    assert abs(p / exp - 1.0) < 1e-3, (p, exp)


# ---------------------------------------------------------------------- IISPH
def test_source_term_vde_matches_the_reference_manual_formula():
    """iisph.rs:746-778: two approaching particles, V = 0.02, dt = 0.05:
    s_0 = -dt sum_j V_j (v*_0 - v*_j) . grad W_0j."""
    dt = 0.05
    x = POS3[:2]
    o = make(x)
    o.set_field("dt", dt)
    o.set_field("V", [0.02, 0.02])
    o.set_field("pred_v", [[1.0, 0.0, 0.0], [-1.0, 0.0, 0.0]])
    o.call("source_vde")
    s = o.field("s")
    exp = -dt * 0.02 * float(np.array([2.0, 0.0, 0.0]) @ grad_w(x[0] - x[1], 1.0))
    assert abs(s[0] - exp) <= 1e-14
    assert s[0] < 0 and s[1] < 0  # approaching: predicted compression


def test_source_term_boundary_part():
    """iisph.rs:117-251 boundary sum: s = -dt V_b (v*_i - v_b) . grad W_ib. A
    particle moving towards a static wall sample is compressing (s < 0); moving
    away, expanding (s > 0)."""
    dt, xb = 0.05, np.array([0.2, 0.0, 0.0])
    o = make([[0.0, 0.0, 0.0]], boundaries=[one_static_sample(xb)])
    o.sample_volumes(0, [0.01])
    o.set_field("dt", dt)
    o.set_field("V", [0.02])
    g = grad_w(-xb, 1.0)
    for u in (1.0, -1.0):
        o.set_field("pred_v", [[u, 0.0, 0.0]])
        o.call("source_vde")
        s = o.field("s")[0]
        assert abs(s - (-dt * 0.01 * u * g[0])) <= 1e-15
        assert s * u < 0


def test_source_term_is_dt_times_the_velocity_divergence():
    """s_i = 1 - V0/V_i - dt sum_j V_j (v_i - v_j).grad W_ij is the predicted
    relative density deficit 1 - rho*/rho0 = dt div(v) for V = V0. A
    converging linear field v = -beta x (div v = -3 beta) on a lattice gives
    s = -3 beta dt at an interior particle; a diverging one the opposite sign."""
    X, c = lattice(4)
    dx, dt, beta = 0.02, 1e-3, 0.5
    o = make(X, m=np.full(len(X), 1000 * dx ** 3), h=4 * dx, dx=dx)
    o.set_field("dt", dt)
    o.set_field("V", np.full(len(X), dx ** 3))
    for sign in (-1.0, 1.0):
        o.set_field("pred_v", sign * beta * X)
        o.call("source_vp")
        s = o.field("s")[c]
        assert abs(s / (sign * 3 * beta * dt) - 1.0) < 1e-3, (sign, s)


def test_source_term_vp_includes_the_volume_deviation():
    """iisph.rs:780-811: an isolated particle at rest with V = V0 / 2 has
    s = 1 - V0/V = -1."""
    dx = 0.3
    o = make([[0.0, 0.0, 0.0]], dx=dx)
    o.set_field("dt", 0.05)
    o.set_field("V", [0.5 * dx ** 3])
    o.set_field("pred_v", [[0.0, 0.0, 0.0]])
    o.call("source_vp")
    assert abs(o.field("s")[0] + 1.0) <= 1e-15


def test_diagonal_element_matches_the_derivation():
    """iisph.rs:254-342, 852-1020: a_ii = dt^2 c_i . (S_f + S_b) + S2 with
    S_f = sum_j V_j grad W_ij, S_b = sum_b V_b grad W_ib,
    c_i = -V_i/m_i (S_f + 2 w S_b), S2 = -dt^2 sum_j V_i V_j^2 / m_j |grad W_ij|^2;
    zero for an isolated particle. Three fluid particles of unequal masses and
    one boundary sample (weighting 0.5)."""
    dt, w, h = 0.05, 0.5, 1.0
    xb = np.array([0.2, -0.1, 0.0])
    m = np.array([0.5, 0.8, 1.1])
    vols = np.array([0.02, 0.025, 0.03])
    o = make(POS3, m=m, boundaries=[one_static_sample(xb)],
             boundary_pressure_acceleration_weighting=w)
    o.sample_volumes(0, [0.01])
    o.set_field("dt", dt)
    o.set_field("V", vols)
    o.call("diagonal")
    a_ii = o.field("a_ii")
    for i in range(3):
        sf, s2 = np.zeros(3), 0.0
        for j in neighbors_of(o, i):
            g = grad_w(POS3[i] - POS3[j], h)
            sf += vols[j] * g
            s2 -= dt ** 2 * vols[i] * vols[j] ** 2 / m[j] * float(g @ g)
        sb = 0.01 * grad_w(POS3[i] - xb, h)
        ci = -vols[i] / m[i] * (sf + 2 * w * sb)
        exp = dt ** 2 * float(ci @ (sf + sb)) + s2
        assert abs(a_ii[i] - exp) <= 1e-13 * abs(exp), i
        assert a_ii[i] < 0
    lone = make([[0.0, 0.0, 0.0]])
    lone.set_field("dt", dt)
    lone.call("diagonal")
    assert lone.field("a_ii")[0] == 0.0


def test_initialize_rules():
    """iisph.rs:344-376, 1025-1080: p = max(omega s / a_ii, 0) where |a_ii| >
    eps, else 0."""
    far = np.array([[0.0, 0, 0], [1000.0, 0, 0], [0, 1000.0, 0]])
    o = make(far, relaxation_factor=0.5, min_diagonal_element=1e-6)
    o.set_field("s", [-5.0, 1.0, -1.0])
    o.set_field("a_ii", [0.0, -2.0, -2.0])
    o.call("initialize")
    assert list(o.field("p")) == [0.0, 0.0, 0.25]


# ----------------------------------------------------------------- IISPHwOST
def test_iisphwost_isolated_particle_hand_derivation():
    """iisph_optimized_source_term.rs:531-586: an isolated particle gets no
    pressure; the acceleration is overwritten to zero and the committed
    state is v + dt g, x + dt (v + dt g) (TakePredicted takes it next step)."""
    dt = 0.05
    x0, v0 = np.array([[1.0, 2.0, 3.0]]), np.array([[0.5, 0.0, 0.0]])
    o = make(x0, v0, dx=0.3, solver="IISPHwOST", integrator="TakePredicted",
             time_increment=dt)
    st = o.state()
    assert st["p"][0] == 0.0 and np.all(st["a"][0] == 0.0)
    o.step(1)
    st = o.state()
    g = np.array([0.0, 0.0, -9.81])
    assert np.allclose(st["v"][0], v0[0] + dt * g, rtol=0, atol=1e-14)
    assert np.allclose(st["x"][0], x0[0] + dt * (v0[0] + dt * g), rtol=0, atol=1e-14)


def test_iisphwost_commit_is_euler_cromer_consistent():
    """IISPHwOST writes x* = x + dt v*_1 + dt^2 a_2 and v* = v*_1 + dt a_2
    (iisph_optimized_source_term.rs:138-196), so the committed step satisfies
    x_{n+1} - x_n = dt v_{n+1} exactly up to rounding, whatever the pressure
    acceleration a_2 of EQS2 is. A compressed block under load checks it with
    a non-zero a_2: the jittered block's first solve (System::new's update)."""
    sc = scenes.small_iisph_box(side=10, solver="IISPHwOST", jitter=2e-3)
    sc.params.cfl = False
    sc.params.time_increment = dt = 2e-3
    o = refcpu.OracleSystem(sc)
    st0 = o.state()
    assert np.max(np.abs(st0["a"])) > 1.0  # EQS2 pressure acceleration present
    o.step(1)
    st1 = o.state()
    dev = np.max(np.abs((st1["x"] - st0["x"]) - dt * st1["v"]))
    assert dev <= 8e-15, dev
    assert np.max(np.abs(dt * dt * st0["a"])) > 1e3 * dev


# ----------------------------------------------------------- volume-map samples
def planar_lattice(value_fn):
    """A single serendipity cell over [-1, 1]^3 (dx = 2) with node values
    value_fn(node) (volume_map_boundary.rs:2921-2945)."""
    n = np.array([1, 1, 1], dtype=np.int64)
    ax = -1.0 + np.arange(4) / 3.0 * 2.0
    z, y, x = np.meshgrid(ax, ax, ax, indexing="ij")
    pts = np.stack([x, y, z], -1)
    return scenes.Lattice(np.array([-1.0, -1.0, -1.0]), np.array([1.0, 1.0, 1.0]), 2.0, n,
                          np.ascontiguousarray(value_fn(pts), dtype=np.float64),
                          np.ones((4, 4, 4), dtype=np.uint8))


@pytest.mark.parametrize("px", [0.1, 0.02, 0.3])
def test_volume_map_sample_on_a_plane(px):
    """volume_map_boundary.rs:1283-1378, 2921-3010: SDF sd = x (reproduced
    exactly by the serendipity element), constant volume map 1; a particle at
    (px, 0, 0), h = 0.5, dx = 0.1 gets its sample at x - (1,0,0) max(px +
    dx/4, dx) with V = min(1, 4/3 pi h^3), and four viscosity samples at
    +-h/2 along t1 = (0,0,-1), t2 = (0,-1,0) with V/4 each."""
    h, dx = 0.5, 0.1
    sdf = planar_lattice(lambda p: p[..., 0])
    vm = planar_lattice(lambda p: np.ones(p.shape[:-1]))
    b = scenes.VolumeMap(sdf=sdf, vm=vm)
    o = make([[px, 0.0, 0.0]], h=h, dx=dx, boundaries=[b])
    normal = o.boundary_samples(0, 0, 0)
    if px >= h:
        assert len(normal) == 0
        return
    assert len(normal) == 1
    d = max(px + 0.25 * dx, dx)
    xb = np.array([px - d, 0.0, 0.0])
    vol = min(1.0, 4.0 / 3.0 * math.pi * h ** 3)
    assert np.allclose(normal[0, :3], xb, rtol=0, atol=1e-14)
    assert abs(normal[0, 6] - vol) <= 1e-15
    visc = o.boundary_samples(0, 0, 1)
    t1, t2 = np.array([0.0, 0.0, -1.0]), np.array([0.0, -1.0, 0.0])
    exp = [xb + 0.5 * h * t1, xb - 0.5 * h * t1, xb + 0.5 * h * t2, xb - 0.5 * h * t2]
    assert len(visc) == 4
    for k in range(4):
        assert np.allclose(visc[k, :3], exp[k], rtol=0, atol=1e-14), k
        assert abs(visc[k, 6] - 0.25 * vol) <= 1e-15


# -------------------------------------------------------------------- rigid body
def test_rigid_torque_is_r_cross_f():
    """rigid_body_motion.rs:143-146: tau += (x - c) x f. A force +y applied at
    +x of the centre turns the body counter-clockwise about +z (right-hand
    rule): tau = (0, 0, |r||f|)."""
    r = refcpu.Rigid(2.0, np.diag([1.0, 2.0, 3.0]), [1.0, 1.0, 1.0])
    r.add_force([0.0, 3.0, 0.0], [3.0, 1.0, 1.0])
    assert np.array_equal(r.torque, [0.0, 0.0, 6.0])
    assert np.array_equal(r.force, [0.0, 3.0, 0.0])
    r.add_force([0.0, 0.0, 1.0], [1.0, 2.0, 1.0])  # +z at +y: tau_x = +1
    assert np.array_equal(r.torque, [1.0, 0.0, 6.0])


def test_rigid_step_integrates_momenta():
    """rigid_body_motion.rs:183-201: v += F/M dt, c += v dt, L += tau dt,
    omega = I_w^-1 L; force and torque reset."""
    r = refcpu.Rigid(2.0, np.diag([1.0, 2.0, 4.0]), [0.0, 0.0, 0.0])
    r.add_force([4.0, 0.0, 0.0], [0.0, 0.0, 0.0])
    r.add_force([0.0, 2.0, 0.0], [1.0, 0.0, 0.0])  # tau = (0, 0, 2)
    r.step(0.5)
    assert np.allclose(r.lin_vel, [1.0, 0.5, 0.0])
    assert np.allclose(r.com, [0.5, 0.25, 0.0])
    assert np.allclose(r.ang_mom, [0.0, 0.0, 1.0])
    assert np.allclose(r.ang_vel, [0.0, 0.0, 0.25])  # I_zz = 4 at the start orientation
    assert np.all(r.force == 0) and np.all(r.torque == 0)


def test_fixed_step_time_is_steps_times_dt():
    """sph/mod.rs:107-116 without cfl_time_step: time = offset + steps * dt, so
    after 10 steps of 1e-3 the time is 10 * 1e-3 = 0.01 exactly (a running sum
    gives 0.010000000000000002)."""
    o = make([[0.0, 0.0, 0.0]], time_increment=1e-3)
    o.step(10)
    t = o.measure()["time"]
    assert t == 10 * 1e-3 and t == 0.01


@pytest.mark.parametrize("com, lin_vel, ang_vel, fluid_at, pos_want, vel_want", [
    # update_positions_and_velocities_transforms_body_frame_to_world
    # (static_sample_boundary.rs:649-658): identity orientation, p + c
    ((5.0, 0.0, 0.0), (0.0, 0.0, 0.0), (0.0, 0.0, 0.0), (6.01, 0.0, 0.0),
     (6.0, 0.0, 0.0), (0.0, 0.0, 0.0)),
    # update_positions_and_velocities_computes_rigid_body_velocity (:660-679):
    # v + omega x (p - c) = (1, 0, 0) + (0, 0, 2) x (1, 0, 0) = (1, 2, 0)
    ((0.0, 0.0, 0.0), (1.0, 0.0, 0.0), (0.0, 0.0, 2.0), (1.01, 0.0, 0.0),
     (1.0, 0.0, 0.0), (1.0, 2.0, 0.0)),
])
def test_moving_sample_pose_and_velocity(com, lin_vel, ang_vel, fluid_at, pos_want, vel_want):
    """A StaticSampleBoundary sample on a rigid body, body-frame (1, 0, 0),
    as a fluid particle within h of it sees it after System::new: the
    reference's own known answers for its world position and velocity."""
    dx = 0.02
    p = scenes.Params(kernel_support_radius=2 * dx, rest_density_grid_spacing=dx,
                      solver="SESPH", integrator="EulerCromer")
    rb = scenes.Rigid(mass=1.0, inertia_body=np.eye(3), com=np.array(com),
                      lin_vel=np.array(lin_vel), ang_vel=np.array(ang_vel))
    x = np.array([fluid_at])
    sc = scenes.Scene("moving_sample", p, x, np.zeros((1, 3)), np.array([dx ** 3 * 1000.0]),
                      np.zeros(1, dtype=np.uint32),
                      [scenes.StaticSamples(np.array([[1.0, 0.0, 0.0]]), rb)])
    o = refcpu.OracleSystem(sc)
    s = o.boundary_samples(0, 0)
    assert s.shape == (1, 7)
    assert tuple(s[0, :3]) == pos_want
    assert tuple(s[0, 3:6]) == vel_want


# --------------------------------------------- other integrators and solvers
def w_kernel(r, h):
    q = math.sqrt(float(np.asarray(r) @ np.asarray(r))) / h
    return ALPHA / h ** 3 * (max(1.0 - q, 0.0) ** 3 - 4.0 * max(0.5 - q, 0.0) ** 3)


def integrate_one(integrator, x, v, a, dt, slot=None):
    o = make([x], v=[v], solver="SESPH", integrator=integrator)
    o.set_field("x", [x])
    o.set_field("v", [v])
    o.set_field("a", [a])
    o.set_field("dt", dt)
    if slot is not None:
        o.set_field("slot_x", [slot])
    o.call("integrate")
    return o.field("x")[0], o.field("v")[0], o.field("slot_x")[0]


def test_explicit_euler_known_answers():
    """tests/sph/integration_schemes/explicit_euler.rs:84-135: x += dt v_old,
    v += dt a (the position takes the velocity from the start of the step)."""
    x0, v0, a0, dt = np.array([1.0, 2.0, 3.0]), np.array([0.5, -0.5, 1.0]), np.array([2.0, 0.0, -9.81]), 0.1
    x, v, _ = integrate_one("ExplicitEuler", x0, v0, a0, dt)
    assert np.allclose(x, x0 + dt * v0, rtol=0, atol=1e-15)
    assert np.allclose(v, v0 + dt * a0, rtol=0, atol=1e-15)
    xe, _, _ = integrate_one("EulerCromer", x0, v0, a0, dt)
    assert not np.allclose(xe, x)  # Euler-Cromer moves with the new velocity


def test_verlet_known_answers():
    """tests/sph/integration_schemes/verlet.rs:106-184: x' = x + (x - x_prev) +
    dt^2 a, v = (x' - x) / dt, and the slot keeps the pre-step position; a
    particle at rest with x_prev = x and no acceleration stays put."""
    xn, xp, a, dt = np.array([1.0, 0.0, 0.0]), np.array([0.5, 0.0, 0.0]), np.array([0.0, -9.81, 0.0]), 0.1
    x, v, slot = integrate_one("Verlet", xn, np.zeros(3), a, dt, slot=xp)
    exp = xn + (xn - xp) + dt * dt * a
    assert np.allclose(x, exp, rtol=0, atol=1e-15)
    assert np.allclose(v, (exp - xn) / dt, rtol=0, atol=1e-13)
    assert np.array_equal(slot, xn)
    x, v, _ = integrate_one("Verlet", xn, np.zeros(3), np.zeros(3), dt, slot=xn)
    assert np.array_equal(x, xn) and np.array_equal(v, np.zeros(3))


def test_sesph_splitting_compressed_pair():
    """tests/sph/pressure_solver/sesph_with_splitting.rs:376-429: two
    particles 0.3 apart (h = 1, m = 0.5, k = 500, dt = 0.05): rho*_i =
    sum_j m_j W_ij (the self pair included), p = k max(V0 / V - 1, 0) with
    V = m / rho*; with the particles approaching (v*_0 = +u e_x,
    v*_1 = -u e_x) the density gains dt (v*_i - v*_j) . grad W_ij. The
    spacing is 0.9 instead of the reference test's 0.3, whose rest volume
    leaves both pressures at zero. With unequal masses (0.5 and 0.8) each
    neighbor term carries the neighbor's mass."""
    h, dx, k, dt = 1.0, 0.9, 500.0, 0.05
    V0 = (dx * dx) * dx
    x = np.array([[0.0, 0.0, 0.0], [0.3, 0.0, 0.0]])
    for ms in ((0.5, 0.5), (0.5, 0.8)):
        m = np.array(ms)
        for u in (0.0, 0.4):
            v = np.array([[u, 0.0, 0.0], [-u, 0.0, 0.0]])
            o = make(x, v=v, m=m, h=h, dx=dx, stiffness=k, solver="SESPHwSplitting")
            o.set_field("v", v)
            o.set_field("a", np.zeros((2, 3)))
            o.set_field("dt", dt)
            o.call("sesph_splitting")
            p = o.field("p")
            for i in range(2):
                r = x[i] - x[1 - i]
                rho = m[i] * w_kernel(np.zeros(3), h) + m[1 - i] * w_kernel(r, h) + \
                    dt * float((v[i] - v[1 - i]) @ grad_w(r, h))
                exp = k * max(V0 / (m[i] / rho) - 1.0, 0.0)
                assert abs(p[i] - exp) <= 1e-9 * max(exp, 1.0), (ms, u, i, p[i], exp)
            if ms[0] == ms[1]:
                mm = ms[0]
                static = k * max(V0 / (mm / (mm * w_kernel(np.zeros(3), h) +
                                             mm * w_kernel(x[1], h))) - 1, 0)
                assert static > 0 and (p[0] > static if u > 0 else p[0] == static)


def test_sesph_splitting_boundary_mirrors_the_particle_mass():
    """sesph_with_splitting.rs:132-148 and tests/...:432-473: a boundary sample
    of volume V_b adds V_b m_i / V0 W_ib + dt v*_i . grad W_ib."""
    h, dx, k, dt, m, Vb = 1.0, 0.9, 500.0, 0.05, 0.5, 0.2
    V0 = (dx * dx) * dx
    xb = np.array([[0.2, 0.0, 0.0]])
    u = np.array([[0.0, 0.0, 0.0]]), np.array([[-0.3, 0.0, 0.0]])
    for v in u:
        o = make([[0.0, 0.0, 0.0]], v=v, m=[m], h=h, dx=dx, stiffness=k, solver="SESPHwSplitting",
                 boundaries=[scenes.StaticSamples(xb)])
        o.sample_volumes(0, [Vb])
        o.set_field("v", v)
        o.set_field("a", np.zeros((1, 3)))
        o.set_field("dt", dt)
        o.call("neighbors")
        o.call("sesph_splitting")
        r = -xb[0]
        rho = m * w_kernel(np.zeros(3), h) + (Vb * m / V0 * w_kernel(r, h) + dt * float(v[0] @ grad_w(r, h)))
        exp = k * max(V0 / (m / rho) - 1.0, 0.0)
        assert exp > 0
        assert abs(o.field("p")[0] - exp) <= 1e-9 * max(exp, 1.0), (v, o.field("p")[0], exp)
