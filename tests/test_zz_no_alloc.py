"""The allocation-free neighbor search (sph_neighbors_find_noalloc) against
the allocating one (sph_neighbors_find), through the C ABI.

Device buffers for the inputs and outputs are allocated here with hipMalloc
from the HIP runtime the library itself links (ctypes); the entry point
allocates nothing. Bar: offsets and indices bit-identical, on two calls in a
row, the second with a smaller input (the static scratch keeps the first
call's contents).
"""
import ctypes as C

import numpy as np
import pytest

from sph_amd import _native as N
from sph_amd import scenes, system

pytestmark = pytest.mark.gpu

MAX_POINTS = 1 << 18
H2D, D2H = 1, 2  # hipMemcpyHostToDevice, hipMemcpyDeviceToHost


@pytest.fixture(scope="module")
def hip():
    N.lib()  # loads libsphhip.so and with it the HIP runtime it links
    rt = C.CDLL("libamdhip64.so.7")
    rt.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    rt.hipFree.argtypes = [C.c_void_p]
    rt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    rt.hipDeviceSynchronize.argtypes = []
    return rt


class DeviceArray:
    """A device copy of a numpy array (or an uninitialised one)."""

    def __init__(self, rt, host=None, nbytes=None):
        self.rt = rt
        self.nbytes = int(host.nbytes if host is not None else nbytes)
        self.p = C.c_void_p()
        assert rt.hipMalloc(C.byref(self.p), max(self.nbytes, 8)) == 0
        if host is not None and self.nbytes:
            h = np.ascontiguousarray(host)
            assert rt.hipMemcpy(self.p, h.ctypes.data, self.nbytes, H2D) == 0

    def get(self, dtype, count):
        out = np.empty(count, dtype=dtype)
        if count:
            assert self.rt.hipMemcpy(out.ctypes.data, self.p, out.nbytes, D2H) == 0
        return out

    def __del__(self):
        self.rt.hipFree(self.p)


def noalloc(rt, sh, search_range, pos, smp, cap=None):
    """One call of the no-alloc entry point on device copies of pos / smp."""
    pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
    smp = np.ascontiguousarray(smp, dtype=np.float64).reshape(-1, 3)
    n, m = len(pos), len(smp)
    if cap is None:
        cap = max(64 * n, 1)
    dpos, dsmp = DeviceArray(rt, pos), DeviceArray(rt, smp)
    doff = DeviceArray(rt, np.full(n + 1, 0xffffffffffffffff, dtype=np.uint64))
    didx = DeviceArray(rt, np.full(max(cap, 1), 0xffffffffffffffff, dtype=np.uint64))
    assert rt.hipDeviceSynchronize() == 0
    total = C.c_size_t(0)
    st = N.lib().sph_neighbors_find_noalloc(sh._ctx.h, sh.cell_size, float(search_range),
                                            dpos.p if n else None, n,
                                            dsmp.p if m else None, m,
                                            doff.p, didx.p, cap, C.byref(total))
    assert rt.hipDeviceSynchronize() == 0
    off = doff.get(np.uint64, n + 1)
    idx = didx.get(np.uint64, min(int(total.value), cap))
    return st, off, idx, int(total.value)


def case_inputs(name):
    rng = np.random.default_rng(11)
    if name == "random_large":  # ~4.6e5 neighbors, cells on both sides of zero
        pos = rng.uniform(-0.3, 0.3, size=(20000, 3))
        smp = rng.uniform(-0.3, 0.3, size=(25000, 3))
        return 0.04, 0.04, pos, smp
    if name == "lattice_boundary":  # fluid -> static samples of the dam break
        sc = scenes.dam_break_sesph()
        h = sc.params.kernel_support_radius
        return h, h, sc.x, sc.boundaries[0].positions
    if name == "wide_range_small":  # range = 2 x cell: 5^3 stencil, smaller than both above
        pos = rng.uniform(0, 1, size=(1500, 3))
        smp = rng.uniform(0, 1, size=(1800, 3))
        return 0.05, 0.1, pos, smp
    raise KeyError(name)


def compare(rt, name):
    cs, rng_, pos, smp = case_inputs(name)
    print(f"no-alloc case {name}: n={len(pos)} m={len(smp)} cell={cs} range={rng_}", flush=True)
    sh = system.SpatialHashing(cs)
    ref = sh.find_samples(rng_, pos, smp)
    st, off, idx, total = noalloc(rt, sh, rng_, pos, smp)
    assert st == 0, N.lib().sph_last_error()
    assert total == len(ref.indices) and total > 0
    assert np.array_equal(off, ref.offsets)
    assert np.array_equal(idx, ref.indices)


def test_two_calls_in_a_row_second_smaller(hip):
    compare(hip, "random_large")
    compare(hip, "lattice_boundary")
    compare(hip, "wide_range_small")


def test_empty_inputs(hip):
    sh = system.SpatialHashing(0.1)
    print("no-alloc case empty inputs", flush=True)
    st, off, idx, total = noalloc(hip, sh, 0.1, np.zeros((0, 3)), np.zeros((5, 3)))
    assert st == 0 and total == 0 and np.array_equal(off, np.zeros(1, np.uint64))
    st, off, idx, total = noalloc(hip, sh, 0.1, np.zeros((4, 3)), np.zeros((0, 3)))
    assert st == 0 and total == 0 and np.array_equal(off, np.zeros(5, np.uint64))


def test_capacity_error_reports_total(hip):
    cs, rng_, pos, smp = case_inputs("wide_range_small")
    print("no-alloc case capacity", flush=True)
    sh = system.SpatialHashing(cs)
    ref = sh.find_samples(rng_, pos, smp)
    st, off, _, total = noalloc(hip, sh, rng_, pos, smp, cap=10)
    assert st == -4 and total == len(ref.indices)
    assert np.array_equal(off, ref.offsets)


def test_sizes_beyond_the_static_limit_are_refused(hip):
    sh = system.SpatialHashing(0.1)
    print("no-alloc case over the limit", flush=True)
    total = C.c_size_t(7)
    # refused before any device access: the pointers are never dereferenced
    buf = DeviceArray(hip, nbytes=64)
    st = N.lib().sph_neighbors_find_noalloc(sh._ctx.h, 0.1, 0.1, buf.p, MAX_POINTS + 1,
                                            buf.p, 3, buf.p, buf.p, 8, C.byref(total))
    assert st == -5 and b"262144" in N.lib().sph_last_error()
