"""CPU pins of the oracle's volume-map construction (ref_vm_* in
oracle/refcpu.c, following volume_map_boundary.rs:1593-1624, 1684-1726 and
utilities/discretization.rs:33-79).

The reference has no tests of this path; the checks are closed forms of the
ball integral (constant and planar signed-distance fields), the early returns
of SDFnWrapper::volume, and a brute-force restatement of the cell rule on a
tiny lattice. The Gauss-Legendre nodes themselves come from the gauss-quad
crate (0.3.2) in the reference; here they are an input (numpy's leggauss), so
the node table is parity unpinned.
"""
import math

import numpy as np
import pytest

import refcpu
from sph_amd import scenes

GX, GW = np.polynomial.legendre.leggauss(30)


def ball(h):
    return 4.0 / 3.0 * math.pi * h ** 3


def full_lattice(fn, lo, hi, dx):
    """Lattice over [lo, hi] with every node present."""
    return scenes._lattice(lo, hi, 0.0, dx, None, None, None, fn)


def plane(normal, p0):
    nrm = np.asarray(normal, dtype=np.float64)
    nrm = nrm / np.linalg.norm(nrm)
    return lambda p: (p - np.asarray(p0)) @ nrm


def test_constant_fields_integrate_to_scaled_ball_volume():
    h = 0.05
    c = np.array([0.2, 0.2, 0.2])
    for sd, frac in ((-0.3, 1.0), (0.5 * h, 0.25), (0.0, 1.0), (1.2 * h, 0.0)):
        lat = full_lattice(lambda p, sd=sd: np.full(p.shape[:-1], sd), [0, 0, 0], [0.4] * 3, 0.05)
        st, v = refcpu.vm_integrate_ball(lat, c, h, GX, GW)
        assert st == 0
        # cubic extension: 1 below 0, w(1/2)/w(0) = 0.125/0.5 at h/2, 0 from h on
        assert abs(v - frac * ball(h)) <= 1e-13 * ball(h), (sd, v)


@pytest.mark.parametrize("normal", [(1, 0, 0), (0, 1, 0), (0, 0, 1), (1, -2, 0.5)])
def test_planar_field_matches_halfspace_closed_form(normal):
    # for a linear SDF the ball integral is int_{-h}^{h} gamma(s0 + s) pi (h^2 - s^2) ds
    h = 0.05
    c = np.array([0.21, 0.19, 0.2])
    errs = []
    for s0 in (-0.9 * h, -0.4 * h, -0.05 * h, 0.3 * h, 0.7 * h, 1.4 * h):
        p0 = c - s0 * np.asarray(normal, dtype=np.float64) / np.linalg.norm(normal)
        lat = full_lattice(plane(normal, p0), [0, 0, 0], [0.4] * 3, 0.05)
        st, v = refcpu.vm_integrate_ball(lat, c, h, GX, GW)
        assert st == 0
        exact = float(scenes.halfspace_volume(np.array([s0]), h)[0])
        errs.append(abs(v - exact) / ball(h))
    # degree 30 on the kinked integrand: well below 1e-3 of the ball volume
    assert max(errs) < 1e-3, errs


def test_node_value_early_returns_and_clamp():
    h = 0.05
    c = np.array([0.2, 0.2, 0.2])
    for sd, want in ((2.5 * h, 0.0), (3.0, 0.0), (-1.5 * h, ball(h)), (-7.0, ball(h))):
        lat = full_lattice(lambda p, sd=sd: np.full(p.shape[:-1], sd), [0, 0, 0], [0.4] * 3, 0.05)
        st, v = refcpu.vm_node_value(lat, c, h, GX, GW)
        assert st == 0 and v == want
    # the center outside the SDF lattice: treated as outside the band, value 0
    lat = full_lattice(lambda p: np.full(p.shape[:-1], -1.0), [0, 0, 0], [0.4] * 3, 0.05)
    st, v = refcpu.vm_node_value(lat, np.array([0.5, 0.2, 0.2]), h, GX, GW)
    assert st == 0 and v == 0.0
    # center inside but the ball leaves the lattice: the quadrature fails
    lat = full_lattice(plane((1, 0, 0), (0.0, 0, 0)), [0, 0, 0], [0.4] * 3, 0.05)
    st, v = refcpu.vm_node_value(lat, np.array([0.02, 0.2, 0.2]), h, GX, GW)
    assert st != 0
    st, v = refcpu.vm_integrate_ball(lat, np.array([0.02, 0.2, 0.2]), h, GX, GW)
    assert st != 0


def test_build_follows_node_values_and_cell_rule():
    # tiny lattice wider than the SDF lattice, so that both failing centers
    # (value 0) and failing quadratures (pruned cells) occur
    h = 0.06
    gx, gw = np.polynomial.legendre.leggauss(8)
    sdf = full_lattice(plane((1, 0.3, 0), (0.06, 0, 0)), [0, 0, 0], [0.4] * 3, 0.08)
    x_min = np.array([-0.05, 0.1, 0.1])
    n = np.array([4, 2, 2])
    dx = 0.1
    vals, pres = refcpu.vm_build(sdf, h, x_min, n, dx, gx, gw)
    pts = scenes._node_positions(x_min, n, dx)
    ok = np.zeros(pts.shape[:-1], dtype=bool)
    val = np.zeros(pts.shape[:-1])
    isnode = scenes._present_from_kept(np.ones((n[2], n[1], n[0]), bool), n).astype(bool)
    for idx in zip(*np.nonzero(isnode)):
        st, v = refcpu.vm_node_value(sdf, pts[idx], h, gx, gw)
        ok[idx] = st == 0
        val[idx] = v if st == 0 else 0.0
    keep = np.all(scenes._cell_node_values(ok.astype(np.uint8), n) == 1, axis=0)
    want_pres = scenes._present_from_kept(keep, n)
    assert np.array_equal(pres, want_pres)
    assert 0 < keep.sum() < keep.size           # some cells pruned, some kept
    assert np.array_equal(vals, np.where(want_pres == 1, val, 0.0))
    assert np.all((vals >= 0) & (vals <= ball(h)))


def test_box_container_faces_match_halfspace_volume():
    # away from edges and corners the container SDF is planar, so the
    # quadrature volume map equals the analytic half-space map there
    dx = 0.02
    h = 2.0 * dx
    lo, hi = np.zeros(3), np.full(3, 0.32)
    sd = lambda q: scenes.sd_box_container(q, lo, hi)
    _, vm_q = scenes.volume_map_quadrature(sd, lo, hi, h, dx, refcpu.vm_build, degree=30)
    _, vm_a = scenes.volume_map_fields(sd, lo, hi, h, dx)
    assert np.array_equal(vm_q.n, vm_a.n)
    pts = scenes._node_positions(vm_q.x_min, vm_q.n, vm_q.dx)
    # nodes near one face whose second-nearest face is over 3h away: their
    # ball and the SDF cells it touches see one plane only
    dist = np.minimum(pts - lo, hi - pts)
    near = np.sort(dist, axis=-1)
    planar = (near[..., 1] > 3 * h) & (vm_q.present == 1) & (vm_a.present == 1)
    assert planar.sum() >= 40
    diff = np.abs(vm_q.values - vm_a.values)[planar] / ball(h)
    assert diff.max() < 1e-3
