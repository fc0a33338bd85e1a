"""CPU pins of the oracle's rigid-rigid contact (contact_* and
ref_contact_solve in oracle/refcpu.c, following
VolumeMapBoundary::step_forward_in_time, volume_map_boundary.rs:56-1102 and
1440-1453).

The reference has no tests of this path. The checks here are:

* closed forms of the contact-cell quadrature for two half-space bodies
  (signed distances linear in z, reproduced exactly by the serendipity
  lattice): with an overlap depth below h/4 both cubic extensions stay on one
  cubic piece, so the degree-8 Gauss-Legendre rule integrates the cell sums
  exactly and the volume, centre, normal sensitivities, torques and diagonal
  elements follow from Beta integrals;
* the solver on cell records: a single cell against its linear closed form,
  several cells against the exact linear complementarity solution found by
  enumerating active sets, and the two friction regimes (sticking removes the
  tangential velocity, sliding saturates the Coulomb cone);
* the one-step relaxed projected Jacobi that the absolute Fischer-Burmeister
  tolerances reduce the solve to when contact volumes are small.

The Gauss-Legendre node table is an input (numpy's leggauss); the reference
takes it from the gauss-quad crate (0.3.2), so the node values stay parity
unpinned.
"""
import itertools
import math

import numpy as np
import pytest

import refcpu
from sph_amd import scenes

MU, RELAX = 0.5, 0.5


# --------------------------------------------------------------------------
# quadrature and diagonal: two half-space bodies
# --------------------------------------------------------------------------
def halfspace_scene(depth, **kw):
    sc = scenes.halfspace_contact(depth, **kw)
    fl, blk = sc.boundaries
    rb = blk.rigid
    return sc, dict(h=sc.params.kernel_support_radius, zt=float(fl.aabb[0][2] + fl.aabb[1][2]),
                    half=float(blk.aabb[1][0]), mass=rb.mass, inertia=rb.inertia_body[0, 0],
                    com=rb.com, dt=sc.params.time_increment)


def beta_sums(depth, h):
    """Per unit area: the volume (Beta(4, 4)) and |f| (Beta(3, 4)) integrals of
    s_a s_b and s'_a s_b over the overlap, s = 16 d^3 / h^3 on the cubic
    piece q >= 1/2 of the cubic extension of 2 phi + h."""
    vol = 256.0 * depth ** 7 / h ** 6 * (math.factorial(3) ** 2 / math.factorial(7))
    f = 768.0 * depth ** 6 / h ** 6 * (math.factorial(2) * math.factorial(3) / math.factorial(6))
    return vol, f


@pytest.mark.parametrize("depth", [0.004, 0.0075])
def test_halfspace_cells_match_beta_integrals(depth):
    sc, g = halfspace_scene(depth)
    o = refcpu.OracleSystem(sc)
    o.step(1)
    cells = o.contact_cells()
    h, half, dt = g["h"], g["half"], g["dt"]
    assert len(cells) > 0
    assert np.all(cells[:, 0] == 0) and np.all(cells[:, 1] == 1)
    # the overlap is the block's footprint x one layer of sub-cells
    area = (2 * half) ** 2
    vol_pa, f_pa = beta_sums(depth, h)
    np.testing.assert_allclose(cells[:, 2].sum(), area * vol_pa, rtol=1e-9)
    # uniform subdivision: equal cells
    nc = len(cells)
    np.testing.assert_allclose(cells[:, 2], area * vol_pa / nc, rtol=1e-9)
    # centre: mid-depth (Beta(4, 4) is symmetric), cell centres in x and y
    np.testing.assert_allclose(cells[:, 5], g["zt"] - 0.5 * depth, rtol=0, atol=1e-12)
    xs = np.unique(np.round(cells[:, 3], 12))
    ys = np.unique(np.round(cells[:, 4], 12))
    assert len(xs) * len(ys) == nc
    np.testing.assert_allclose(xs.mean(), g["com"][0], atol=1e-12)
    np.testing.assert_allclose(ys.mean(), g["com"][1], atol=1e-12)
    # normal sensitivities: -|f| z on the floor, +|f| z on the block
    fz = area * f_pa / nc
    np.testing.assert_allclose(cells[:, 8], -fz, rtol=1e-9)
    np.testing.assert_allclose(cells[:, 14], fz, rtol=1e-9)
    np.testing.assert_allclose(cells[:, [6, 7, 12, 13]], 0.0, atol=1e-12 * fz)
    # torques: none on the static floor; (centre - com) x f_b on the block
    np.testing.assert_array_equal(cells[:, 9:12], 0.0)
    r = cells[:, 3:6] - g["com"]
    tb = np.cross(r, cells[:, 12:15])
    np.testing.assert_allclose(cells[:, 15:18], tb, rtol=1e-9, atol=1e-12 * fz)
    # diagonal elements: dt^2 (|f|^2 / m + tau.I^-1 tau) and, with T = diag(1, 1, 0)
    # and an isotropic inertia k, dt/3 (2/m + (r_x^2 + r_y^2 + 2 r_z^2) / k)
    acc_p = dt * dt * (fz * fz / g["mass"] + (tb * tb).sum(axis=1) / g["inertia"])
    np.testing.assert_allclose(cells[:, 18], acc_p, rtol=1e-9)
    acc_f = dt / 3.0 * (2.0 / g["mass"] +
                        (r[:, 0] ** 2 + r[:, 1] ** 2 + 2 * r[:, 2] ** 2) / g["inertia"])
    np.testing.assert_allclose(cells[:, 19], acc_f, rtol=1e-9)


def test_halfspace_small_volumes_take_one_relaxed_jacobi_step():
    """Cell volumes far below the absolute 1e-6 tolerance: the solve stops after
    the initial projected Jacobi step, p_c = max(0, -0.5 / N * V_err / a_cc),
    with N active cells on the block and V_err from the free motion."""
    depth = 0.006
    sc, g = halfspace_scene(depth)
    o = refcpu.OracleSystem(sc)
    F = o.body_state(1)["force"].copy()
    o.step(1)
    ncell, iters = o.contact_stats()
    cells = o.contact_cells()
    assert ncell == len(cells) > 0 and iters == 0
    dt = g["dt"]
    v_free = dt / g["mass"] * F
    verr = -cells[:, 2] + dt * (cells[:, 12:15] @ v_free)
    assert np.all(verr < 0)
    p1 = np.maximum(0.0, -(RELAX / ncell) * verr / cells[:, 18])
    np.testing.assert_allclose(cells[:, 20], p1, rtol=1e-12)
    np.testing.assert_array_equal(cells[:, 21:24], 0.0)  # a static partner: no friction


def test_no_contact_between_a_container_and_an_inner_body():
    """The container's solid lies outside its own AABB, so the overlap with a
    body inside holds no point where both cubic extensions are positive."""
    sc = scenes.small_rigid_scene()
    o = refcpu.OracleSystem(sc)
    o.step(2)
    assert o.contact_stats() == (0, 0)


def test_contact_off_and_static_pairs():
    sc = scenes.contact_stack(sphere=False)
    off = scenes.contact_stack(sphere=False)
    off.contact_degree = 0
    a, b = refcpu.OracleSystem(sc), refcpu.OracleSystem(off)
    a.step(1)
    b.step(1)
    assert a.contact_stats()[0] > 0 and b.contact_stats() == (0, 0)
    cells = a.contact_cells()
    # pairs in (i < j) order with one dynamic member: never container-slab
    assert set(map(tuple, cells[:, :2].astype(int))) <= {(1, 2), (0, 2)}
    # the committed contact force moves the cube: same fluid, different body
    assert not np.array_equal(a.body_state(2)["lin_vel"], b.body_state(2)["lin_vel"])


# --------------------------------------------------------------------------
# the solver on cell records
# --------------------------------------------------------------------------
def body(dynamic, mass=1.0, com=(0, 0, 0), v=(0, 0, 0), L=(0, 0, 0), F=(0, 0, 0),
         T=(0, 0, 0), Iinv=None):
    Iinv = np.eye(3) if Iinv is None else np.asarray(Iinv)
    return np.concatenate([[1.0 if dynamic else 0.0, mass], com, v, L, F, T, Iinv.ravel()])


def cell(a, b, vol, center, fa, ta, fb, tb, bodies, dt):
    """A record with its diagonal elements from the bodies (compute_diagonal)."""
    rec = np.zeros(refcpu.CONTACT_REC)
    rec[0], rec[1], rec[2] = a, b, vol
    rec[3:6], rec[6:9], rec[9:12], rec[12:15], rec[15:18] = center, fa, ta, fb, tb
    n = np.asarray(fa) / max(np.linalg.norm(fa), 1e-15)
    T = np.eye(3) - np.outer(n, n)
    acc_p, S = 0.0, np.zeros((3, 3))
    for idx, f, t in ((a, fa, ta), (b, fb, tb)):
        bd = bodies[idx]
        if bd[0] == 0.0:
            continue
        Ii = bd[17:26].reshape(3, 3)
        acc_p += np.dot(f, f) / bd[1] + np.dot(t, Ii @ t)
        rx = np.cross(np.eye(3), np.asarray(center) - bd[2:5])  # rows e_i x r
        X = -rx  # [r]x
        S += np.eye(3) / bd[1] - X @ Ii @ X
    rec[18] = dt * dt * acc_p
    rec[19] = np.trace(dt * T @ S) / 3.0
    return rec


def trial(bodies, extra, dt):
    out = []
    for i, bd in enumerate(bodies):
        if bd[0] == 0.0:
            out.append((np.zeros(3), np.zeros(3)))
            continue
        v = bd[5:8] + dt / bd[1] * (bd[11:14] + extra[i, :3])
        w = bd[17:26].reshape(3, 3) @ (bd[8:11] + dt * (bd[14:17] + extra[i, 3:]))
        out.append((v, w))
    return out


def test_single_cell_reaches_its_linear_solution():
    dt = 1e-2
    bodies = np.array([body(False), body(True, mass=2.0, com=(0, 0, 0.3), v=(0.1, 0, -1.0),
                                        F=(0, 0, -19.62), Iinv=np.diag([2.0, 3.0, 4.0]))])
    c = cell(0, 1, 0.02, (0.05, -0.02, 0.0), (0, 0, -1.0), (0, 0, 0), (0, 0, 1.0),
             np.cross(np.array([0.05, -0.02, -0.3]), [0, 0, 1.0]), bodies, dt)
    out, extra, it = refcpu.contact_solve(bodies, [c], dt)
    (vb, wb) = trial(bodies, np.zeros((2, 6)), dt)[1]
    verr0 = -c[2] + dt * (vb @ c[12:15] + wb @ c[15:18])
    p_star = max(0.0, -verr0 / c[18])
    assert it > 0
    # stopped once the volume error (p - p*) a_cc is below the 1e-6 tolerance
    assert abs(out[0, 20] - p_star) * c[18] < 1e-6
    np.testing.assert_allclose(out[0, 20], p_star, rtol=1e-4)
    np.testing.assert_allclose(extra[1, :3], out[0, 20] * c[12:15], rtol=1e-12)
    np.testing.assert_allclose(extra[1, 3:], out[0, 20] * c[15:18], rtol=1e-12)
    # the static body's force is accumulated (and never committed); no torque
    np.testing.assert_allclose(extra[0, :3], out[0, 20] * c[6:9], rtol=1e-12)
    np.testing.assert_array_equal(extra[0, 3:], 0.0)


def lcp_by_enumeration(W, q):
    n = len(q)
    for k in range(n + 1):
        for S in itertools.combinations(range(n), k):
            S = list(S)
            p = np.zeros(n)
            if S:
                p[S] = np.linalg.solve(W[np.ix_(S, S)], -q[S])
            w = q + W @ p
            if np.all(p >= -1e-12) and np.all(w >= -1e-12):
                return p
    raise AssertionError("no complementary solution")


@pytest.mark.parametrize("seed", [0, 1, 2, 3])
def test_frictionless_cells_match_enumerated_lcp(seed):
    """One dynamic body on a static one (no tangential velocity, so no
    friction): the pressures solve the linear complementarity problem
    V_err = q + W p >= 0, p >= 0, p . V_err = 0 exactly enumerated."""
    rng = np.random.default_rng(seed)
    dt = 1e-2
    n = 4
    Iinv = np.diag(rng.uniform(1.0, 3.0, 3))
    bodies = np.array([body(False), body(True, mass=1.5, com=(0, 0, 0.2),
                                         v=rng.uniform(-0.5, 0.5, 3), F=(0, 0, -14.7),
                                         L=rng.uniform(-0.1, 0.1, 3), Iinv=Iinv)])
    cells, J = [], []
    for k in range(n):
        center = np.array([rng.uniform(-0.1, 0.1), rng.uniform(-0.1, 0.1), 0.0])
        nrm = np.array([rng.uniform(-0.3, 0.3), rng.uniform(-0.3, 0.3), 1.0])
        fb = nrm / np.linalg.norm(nrm) * rng.uniform(0.5, 1.5)
        tb = np.cross(center - bodies[1, 2:5], fb)
        vol = rng.uniform(0.5, 2.0) * (1 if k % 2 == 0 else -0.2)
        cells.append(cell(0, 1, vol, center, -fb, np.zeros(3), fb, tb, bodies, dt))
        J.append(np.concatenate([fb, tb]))
    J = np.array(J)
    Minv = np.zeros((6, 6))
    Minv[:3, :3] = np.eye(3) / bodies[1, 1]
    Minv[3:, 3:] = Iinv
    W = dt * dt * J @ Minv @ J.T
    vb, wb = trial(bodies, np.zeros((2, 6)), dt)[1]
    q = np.array([-c[2] for c in cells]) + dt * (J @ np.concatenate([vb, wb]))
    p_exact = lcp_by_enumeration(W, q)
    out, extra, it = refcpu.contact_solve(bodies, cells, dt)
    assert np.all(out[:, 20] >= 0.0)
    # the solve stops at a mean Fischer-Burmeister value of 1e-6 (volume units)
    np.testing.assert_allclose(W @ out[:, 20], W @ p_exact, rtol=0, atol=n * 1e-6)
    assert np.abs(W @ p_exact).max() > 0.1
    np.testing.assert_allclose(extra[1], J.T @ out[:, 20], rtol=1e-12, atol=1e-15)
    np.testing.assert_array_equal(out[:, 21:24], 0.0)


def two_body_friction_case(slide):
    """Body 0 sliding at `slide` along x over body 1 pushed into it: one
    contact cell with normal z."""
    dt = 1e-2
    bodies = np.array([
        body(True, mass=1.0, com=(0, 0, 0.2), v=(slide, 0, -0.5), F=(0, 0, -9.81)),
        body(True, mass=2.0, com=(0, 0, -0.2), v=(0, 0, 0.0), F=(0, 0, -19.62))])
    center = np.array([0.0, 0.0, 0.0])
    fa = np.array([0.0, 0.0, 1.0])
    fb = -fa
    c = cell(0, 1, 0.05, center, fa, np.cross(center - bodies[0, 2:5], fa), fb,
             np.cross(center - bodies[1, 2:5], fb), bodies, dt)
    out, extra, it = refcpu.contact_solve(bodies, [c], dt)
    (va, wa), (vb, wb) = trial(bodies, extra, dt)
    n = fa
    T = np.eye(3) - np.outer(n, n)
    vrel = va + np.cross(wa, center - bodies[0, 2:5]) - vb - np.cross(wb, center - bodies[1, 2:5])
    return out[0], T @ vrel, c


def test_friction_sticks_below_the_cone():
    rec, vt, c = two_body_friction_case(slide=0.05)
    lam, p = rec[21:24], rec[20]
    assert p > 0
    assert np.linalg.norm(lam) < MU * p * np.linalg.norm(c[6:9])
    assert np.linalg.norm(vt) < 1e-3  # the friction tolerance of the solve
    assert lam[0] < 0  # opposes the sliding of body 0 along +x


def test_friction_slides_on_the_cone():
    rec, vt, c = two_body_friction_case(slide=50.0)
    lam, p = rec[21:24], rec[20]
    assert p > 0
    # on the cone of the previous iterate's pressure (projected_jacobi bounds
    # lambda with p before its update), which the final p matches to the
    # solve's tolerance
    np.testing.assert_allclose(np.linalg.norm(lam), MU * p * np.linalg.norm(c[6:9]), rtol=1e-4)
    assert lam[0] < 0 and vt[0] > 0  # still sliding, against the friction
