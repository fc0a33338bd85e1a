"""Boundary fields from triangle meshes on the device against the oracle
(DiscretizedBoundaryFields::new, volume_map_boundary.rs:1547-1636;
TriangleMeshWrapper::signed_distance, :1655-1671).

`sph_mesh_sdf_build` (k_mesh_sdf + k_sdf_cells) equals the oracle's
ref_mesh_sdf_build bit for bit: node values and present flags, for solid and
container windings of a box and a geodesic sphere. The full field pipeline
(mesh distance lattice, then the volume-map quadrature from it) matches too,
and a system on device-built mesh fields steps bit-identically to the oracle
on the same fields.
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


def oracle_sdf(v, t, x_min, n, dx, prune):
    return refcpu.Mesh(v, t).sdf_lattice(x_min, n, dx, prune)


MESHES = {
    "box_container": lambda: scenes.box_mesh([0, 0, 0], [0.6, 0.5, 0.4], inward=True),
    "cube": lambda: scenes.box_mesh([-0.1, -0.1, -0.1], [0.1, 0.1, 0.1]),
    "sphere_container": lambda: scenes.icosphere(0.3, subdivisions=3, inward=True),
    "sphere": lambda: scenes.icosphere(0.15, subdivisions=2, center=(0.01, -0.02, 0.03)),
}


@pytest.mark.parametrize("case", list(MESHES))
def test_mesh_sdf_lattice_matches_oracle(native, case):
    v, t = MESHES[case]()
    h, dx = 0.04, 0.02
    pad = 3.1 * h
    x_min, x_max, n = scenes.lattice_grid(v.min(axis=0), v.max(axis=0), pad, 4 * dx)
    print(f"mesh sdf {case}: {len(t)} triangles, cells {n.tolist()}", flush=True)
    gv, gp = native.build_mesh_sdf(v, t, x_min, n, 4 * dx, pad)
    ov, op = oracle_sdf(v, t, x_min, n, 4 * dx, pad)
    np.testing.assert_array_equal(gp, op)
    np.testing.assert_array_equal(bits(gv), bits(ov))
    assert 0 < gp.sum() < gp.size


def test_mesh_fields_pipeline_matches_oracle(native):
    v, t = scenes.icosphere(0.3, subdivisions=2, inward=True)
    h, dx = 0.04, 0.02
    print(f"mesh fields: {len(t)} triangles", flush=True)
    g_sdf, g_vm = scenes.mesh_fields(v, t, h, dx, native.build_mesh_sdf, native.build_volume_map,
                                     degree=8)
    o_sdf, o_vm = scenes.mesh_fields(v, t, h, dx, oracle_sdf, refcpu.vm_build, degree=8)
    for g, o in ((g_sdf, o_sdf), (g_vm, o_vm)):
        np.testing.assert_array_equal(g.present, o.present)
        np.testing.assert_array_equal(bits(g.values), bits(o.values))


def test_system_on_mesh_fields_matches_oracle(native):
    sc = scenes.small_iisph_box(side=8, jitter=2e-3)
    v, t = scenes.box_mesh([0, 0, 0], [0.6, 0.6, 0.6], inward=True)
    h, dx = sc.params.kernel_support_radius, sc.params.rest_density_grid_spacing
    sdf, vm = scenes.mesh_fields(v, t, h, dx, native.build_mesh_sdf, native.build_volume_map,
                                 degree=8)
    sc.boundaries = [scenes.VolumeMap(sdf, vm, None, "box",
                                      scenes.aabb_of(np.zeros(3), np.full(3, 0.6)))]
    print(f"system on mesh fields: {sc.n} particles", flush=True)
    g, o = native.System(sc), refcpu.OracleSystem(sc)
    g.step_forward_in_time(3)
    o.step(3)
    gs, os_ = g.state(), o.state()
    for k in ("x", "v", "a", "V", "p"):
        np.testing.assert_array_equal(bits(gs[k]), bits(os_[k]), err_msg=k)
