"""Volume-map construction on the device against the oracle.

sph_volume_map_build evaluates, per lattice node, the same binary64
expressions in the same order as the oracle's ref_vm_build (the phi, theta and
r sums of the Gauss-Legendre ball rule are folded in node order), with the
sine and cosine tables taken from the host's libm, so values and present flags
must be bit-identical.
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


def same_bits(a, b):
    return np.array_equal(np.ascontiguousarray(a).view(np.uint64),
                          np.ascontiguousarray(b).view(np.uint64))


def check_pair(native, sd_fn, lo, hi, h, dx, degree, system=None):
    sg, sr = {}, {}
    build = lambda *a: native.build_volume_map(*a, system=system, stats=sg)
    sdf, vm_g = scenes.volume_map_quadrature(sd_fn, lo, hi, h, dx, build, degree=degree)
    ref = lambda *a: refcpu.vm_build(*a, stats=sr)
    _, vm_r = scenes.volume_map_quadrature(sd_fn, lo, hi, h, dx, ref, degree=degree)
    print(f"volume map {vm_g.n.tolist()} cells, degree {degree}: "
          f"{int(vm_g.present.sum())} present nodes, {sg['quadrature_nodes']} quadratures",
          flush=True)
    assert sg["quadrature_nodes"] == sr["quadrature_nodes"] > 0
    assert np.array_equal(vm_g.present, vm_r.present)
    assert same_bits(vm_g.values, vm_r.values)
    return vm_g


def test_box_container_bit_identical(native):
    lo, hi = np.zeros(3), np.full(3, 0.32)
    vm = check_pair(native, lambda q: scenes.sd_box_container(q, lo, hi), lo, hi, 0.04, 0.02, 30)
    ball = 4.0 / 3.0 * np.pi * 0.04 ** 3
    inside = vm.values[vm.present == 1]
    assert inside.max() == ball and inside.min() == 0.0
    assert np.any((inside > 0) & (inside < ball))


def test_sphere_container_on_a_system_stream(native):
    c, r = np.full(3, 0.2), 0.2
    sc = scenes.small_iisph_box()
    sim = native.System(sc)
    check_pair(native, lambda q: scenes.sd_sphere_container(q, c, r), c - r, c + r, 0.04, 0.02,
               30, system=sim)
    sim._ctx.close()


def test_failing_quadratures_prune_cells(native):
    # a volume-map lattice wider than its SDF lattice: centers outside give 0,
    # balls leaving the SDF lattice fail and their cells are pruned
    h = 0.06
    gx, gw = np.polynomial.legendre.leggauss(8)
    nrm = np.array([1.0, 0.3, 0.0]) / np.linalg.norm([1.0, 0.3, 0.0])
    sdf = scenes._lattice([0, 0, 0], [0.4] * 3, 0.0, 0.08, None, None, None,
                          lambda p: (p - np.array([0.06, 0, 0])) @ nrm)
    x_min, n, dx = np.array([-0.05, 0.1, 0.1]), np.array([4, 2, 2]), 0.1
    vg, pg = native.build_volume_map(sdf, h, x_min, n, dx, gx, gw)
    vr, pr = refcpu.vm_build(sdf, h, x_min, n, dx, gx, gw)
    assert np.array_equal(pg, pr) and same_bits(vg, vr)
    assert 0 < pg.sum() < pg.size


def test_rejects_bad_degree(native):
    lo, hi = np.zeros(3), np.full(3, 0.1)
    sdf = scenes._lattice(lo, hi, 0.0, 0.05, None, None, None, lambda p: p[..., 0] - 0.05)
    for deg in (0, 33):
        gx = np.zeros(deg)
        with pytest.raises(Exception, match="degree"):
            native.build_volume_map(sdf, 0.02, lo, np.array([1, 1, 1]), 0.1, gx, gx)
