"""CPU pins of the oracle's signed distance to a triangle mesh (ref_mesh_* in
oracle/refcpu.c, TriangleMeshWrapper::signed_distance,
volume_map_boundary.rs:1655-1671) and of the signed-distance lattice built
from it (DiscretizedBoundaryFields::new, :1547-1591).

The reference evaluates parry3d-f64 0.28's oriented TriMesh; that crate
cannot run here, so its closest-feature tie-breaking and the bits of its
projection stay parity unpinned. The checks:

* the reference's own known answers on its cube fixture (outside +1, inside
  -1, on the surface 0; tests/golden/reference_known_answers.json);
* closed forms: a box mesh gives the exact signed distance of the box at
  random points inside, outside, near edges and corners, in both windings
  (solid and container);
* for a convex geodesic sphere: inside, the distance is the largest signed
  face-plane distance; outside, it is positive, at least that, and at most
  the distance to the nearest vertex; just off vertices and edges (closest
  feature a vertex or an edge, signed by its pseudo-normal) the sign follows
  the side;
* the lattice's pruning rule against the numpy restatement used for the
  analytic scene lattices, on a box container.
"""
import json
import os

import numpy as np
import pytest

import refcpu
from sph_amd import scenes

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "reference_known_answers.json")


def test_reference_cube_known_answers():
    d = json.load(open(GOLDEN))["mesh_signed_distance"]
    m = refcpu.Mesh(np.array(d["vertices"], dtype=np.float64), d["triangles"])
    got = m.signed_distance(np.array(d["points"], dtype=np.float64))
    np.testing.assert_allclose(got, d["expected"], rtol=0, atol=d["tolerance"])


@pytest.mark.parametrize("inward", [False, True])
def test_box_mesh_is_the_box_signed_distance(inward):
    rng = np.random.default_rng(4)
    lo, hi = np.array([0.1, -0.3, 0.2]), np.array([0.9, 0.5, 1.4])
    m = refcpu.Mesh(*scenes.box_mesh(lo, hi, inward))
    assert m.edges == 18
    c, half = (lo + hi) / 2, (hi - lo) / 2
    pts = [c + rng.uniform(-2, 2, (400, 3)) * half,            # everywhere
           c + rng.uniform(-1, 1, (200, 3)) * half * 0.999]     # inside
    # near edges and corners: a corner plus small offsets on either side
    corners = np.array([[lo[0], lo[1], lo[2]], [hi[0], lo[1], hi[2]], [hi[0], hi[1], lo[2]]])
    pts.append((corners[:, None, :] + rng.uniform(-0.05, 0.05, (3, 100, 3))).reshape(-1, 3))
    pts = np.concatenate(pts)
    sd = scenes.sd_box_container(pts, lo, hi) if inward else scenes.sd_box_solid(pts, lo, hi)
    np.testing.assert_allclose(m.signed_distance(pts), sd, rtol=1e-12, atol=1e-14)


def sphere_planes(v, t):
    a, b, c = v[t[:, 0]], v[t[:, 1]], v[t[:, 2]]
    n = np.cross(b - a, c - a)
    n /= np.linalg.norm(n, axis=1)[:, None]
    return n, a


def test_convex_sphere_mesh_inside_is_the_largest_plane_distance():
    v, t = scenes.icosphere(0.7, subdivisions=2, center=(0.1, 0.2, -0.1))
    m = refcpu.Mesh(v, t)
    n, a = sphere_planes(v, t)
    rng = np.random.default_rng(5)
    d = rng.normal(size=(500, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    inside = np.array([0.1, 0.2, -0.1]) + d * rng.uniform(0, 0.62, (500, 1))
    planes = ((inside[:, None, :] - a[None]) * n[None]).sum(-1).max(axis=1)
    assert np.all(planes < 0)
    np.testing.assert_allclose(m.signed_distance(inside), planes, rtol=1e-12, atol=1e-14)
    outside = np.array([0.1, 0.2, -0.1]) + d * rng.uniform(0.75, 1.5, (500, 1))
    sd = m.signed_distance(outside)
    plane_out = ((outside[:, None, :] - a[None]) * n[None]).sum(-1).max(axis=1)
    nearest_vertex = np.linalg.norm(outside[:, None, :] - v[None], axis=-1).min(axis=1)
    assert np.all(sd > 0)
    assert np.all(sd >= plane_out - 1e-14)
    assert np.all(sd <= nearest_vertex + 1e-14)


def test_sign_just_off_vertices_and_edges():
    v, t = scenes.icosphere(1.0, subdivisions=1)
    m = refcpu.Mesh(v, t)
    e_mid = 0.5 * (v[t[:, 0]] + v[t[:, 1]])
    for base in (v, e_mid):
        out = m.signed_distance(base * (1 + 1e-6))
        inn = m.signed_distance(base * (1 - 1e-6))
        assert np.all(out > 0) and np.all(inn < 0)
    # the container winding flips every sign
    mi = refcpu.Mesh(*scenes.icosphere(1.0, subdivisions=1, inward=True))
    p = v * (1 + 1e-3)
    np.testing.assert_array_equal(mi.signed_distance(p), -m.signed_distance(p))


def test_box_container_lattice_matches_the_analytic_lattice():
    h, dx = 0.04, 0.02
    lo, hi = np.zeros(3), np.array([0.6, 0.5, 0.4])
    m = refcpu.Mesh(*scenes.box_mesh(lo, hi, inward=True))
    pad = 3.1 * h
    ref = scenes._lattice(lo, hi, pad, 4 * dx, None, -pad, pad,
                          lambda q: scenes.sd_box_container(q, lo, hi))
    vals, pres = m.sdf_lattice(ref.x_min, ref.n, ref.dx, pad)
    np.testing.assert_array_equal(pres, ref.present)
    np.testing.assert_allclose(vals, ref.values, rtol=1e-12, atol=1e-14)
    assert 0 < pres.sum() < pres.size
