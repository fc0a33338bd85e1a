"""Shared spatial-hash buckets: the reference's duplicate neighbors, on the GPU.

The reference keeps its samples in a HashMap keyed by the 64-bit spatial hash
of their cell (spatial_hashing.rs:27-103). Cells (16, 1135, 349) and
(17, 1133, 348) have equal keys and lie in one 3x3x3 stencil (that of cell
(16, 1134, 348), for one), so a particle there visits their shared bucket
twice and lists the bucket's samples in range twice, merged in sample-index
order at each visit. With h = 0.04 the cells sit near (0.66, 45.4, 13.9) m.
The oracle keeps the reference's hash map, so it reproduces the duplicates;
the HIP path walks such stencils cell by cell through the bucket partners of
each cell and must list and sum exactly what the oracle does.
"""
import copy

import numpy as np
import pytest

import refcpu
from sph_amd import scenes

pytestmark = pytest.mark.gpu

CS = 0.04
# cells (5..9, 5..9, 0..5) of the 0.6 m test boxes moved onto the colliding pair
SHIFT = np.array([10.0, 1127.0, 347.0]) * CS


@pytest.fixture(scope="module")
def native():
    from sph_amd import system
    return system


def shifted(sc, d=SHIFT):
    """The scene moved by d: particles, lattices and static samples."""
    sc = copy.deepcopy(sc)
    sc.x = sc.x + d
    for b in sc.boundaries:
        if isinstance(b, scenes.VolumeMap):
            for lat in (b.sdf, b.vm):
                lat.x_min = lat.x_min + d
                lat.x_max = lat.x_max + d
            if b.aabb is not None:
                b.aabb = (np.asarray(b.aabb[0]) + d, b.aabb[1])
        else:
            b.positions = b.positions + d
    return sc


def duplicates(off, idx):
    return sum(int(off[i + 1] - off[i]) - len(set(idx[off[i]:off[i + 1]].tolist()))
               for i in range(len(off) - 1))


def test_oracle_hash_keys_collide():
    """The two cells' keys are equal (the premise of every test here)."""
    lib = refcpu.lib()
    assert lib.ref_hash(16, 1135, 349) == lib.ref_hash(17, 1133, 348)


def test_neighbors_with_shared_buckets(native):
    rng = np.random.default_rng(3)
    c = np.array([16.5, 1134.0, 348.5]) * CS
    pos = c + rng.uniform(-2.5 * CS, 2.5 * CS, size=(600, 3))
    smp = c + rng.uniform(-2.5 * CS, 2.5 * CS, size=(700, 3))
    print("shared buckets: sph_neighbors_find", flush=True)
    for p, s in ((pos, pos), (pos, smp)):
        nl = native.SpatialHashing(CS).find_samples(CS, p, s)
        off, idx = refcpu.find_samples(CS, CS, p, s)
        assert duplicates(off, idx) > 0
        assert np.array_equal(nl.offsets, off)
        assert np.array_equal(nl.indices, idx)


@pytest.mark.parametrize("case", ["iisph_box", "ost_box", "sesph_static"])
def test_trajectories_with_shared_buckets(native, case):
    """Fluid neighbor lists with duplicates (the one-sweep pair cache) under
    IISPH, IISPHwOST and SESPH with static samples: bit-identical to the
    oracle."""
    base = {"iisph_box": lambda: scenes.small_iisph_box(side=10, jitter=2e-3),
            "ost_box": lambda: scenes.small_iisph_box(side=10, solver="IISPHwOST", jitter=2e-3),
            "sesph_static": lambda: scenes.dam_break_sesph(side=8)}[case]()
    # the dam break's block fills cells (0..4)^3 of its box
    sc = shifted(base, np.array([14.0, 1131.0, 346.0]) * CS if case == "sesph_static" else SHIFT)
    print("shared buckets:", case, flush=True)
    o, g = refcpu.OracleSystem(sc), native.System(sc)
    off, idx = o.fluid_neighbors()
    assert duplicates(off, idx) > 0
    for _ in range(5):
        o.step(1)
        g.step_forward_in_time(1)
        so, sg = o.state(), g.state()
        for k in ("x", "v", "a", "V", "p"):
            assert np.array_equal(sg[k], so[k]), k
        mg, mo = g.take_measurement(), o.measure()
        assert mg["solver_iterations"] == mo["solver_iterations"]


def test_separate_passes_with_shared_buckets(native):
    """The count/fill/volume passes (SPH_OPT_SWEEP = 0) list the same
    duplicates as the one sweep."""
    from sph_amd import _native as N
    sc = shifted(scenes.small_iisph_box(side=10, jitter=2e-3))
    print("shared buckets: separate passes", flush=True)
    a, b = native.System(sc), native.System(sc)
    b.set_option(N.OPT_SWEEP, 0)
    for _ in range(4):
        a.step_forward_in_time(1)
        b.step_forward_in_time(1)
        sa, sb = a.state(), b.state()
        for k in ("x", "v", "a", "V", "p"):
            assert np.array_equal(sa[k], sb[k]), k


def test_quantities_with_shared_buckets(native):
    sc = shifted(scenes.small_iisph_box(side=10, jitter=2e-3))
    print("shared buckets: sensor queries", flush=True)
    o, g = refcpu.OracleSystem(sc), native.System(sc)
    o.step(2)
    g.step_forward_in_time(2)
    rng = np.random.default_rng(5)
    lo, hi = sc.x.min(0) - 0.03, sc.x.max(0) + 0.03
    q = rng.uniform(lo, hi, size=(2000, 3)).astype(np.float32)
    for k, name in enumerate(["Speed", "Volume", "Density", "DensityError", "Pressure",
                              "KineticEnergy"]):
        a = g.get_quantity_at_positions(name, q)
        b = o.quantity_at_positions(k, q)
        assert np.array_equal(np.asarray(a).view(np.uint32), np.asarray(b).view(np.uint32)), name
