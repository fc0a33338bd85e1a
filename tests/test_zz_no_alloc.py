"""The allocation-free neighbor search (sph_neighbors_find_noalloc) against
the allocating one (sph_neighbors_find), through the C ABI.

Device buffers for inputs and outputs come from torch (plumbing only); the
entry point itself allocates nothing. Bar: offsets and indices bit-identical,
on two calls in a row, the second with a smaller input (the static scratch
keeps the first call's contents).
"""
import ctypes as C

import numpy as np
import pytest

from sph_amd import _native as N
from sph_amd import scenes, system

pytestmark = pytest.mark.gpu

MAX_POINTS = 1 << 18


@pytest.fixture(scope="module")
def torch_cuda():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def noalloc(torch, sh, search_range, pos, smp, cap=None):
    """One call of the no-alloc entry point on device copies of pos / smp."""
    pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
    smp = np.ascontiguousarray(smp, dtype=np.float64).reshape(-1, 3)
    n, m = len(pos), len(smp)
    dpos = torch.from_numpy(pos).to("cuda")
    dsmp = torch.from_numpy(smp).to("cuda")
    if cap is None:
        cap = max(64 * n, 1)
    doff = torch.full((n + 1,), -1, dtype=torch.int64, device="cuda")
    didx = torch.full((max(cap, 1),), -1, dtype=torch.int64, device="cuda")
    torch.cuda.synchronize()
    total = C.c_size_t(0)
    st = N.lib().sph_neighbors_find_noalloc(sh._ctx.h, sh.cell_size, float(search_range),
                                            dpos.data_ptr() if n else None, n,
                                            dsmp.data_ptr() if m else None, m,
                                            doff.data_ptr(), didx.data_ptr(), cap,
                                            C.byref(total))
    torch.cuda.synchronize()
    off = doff.cpu().numpy().view(np.uint64)
    idx = didx.cpu().numpy().view(np.uint64)[:int(total.value)]
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


def compare(torch, name):
    cs, rng_, pos, smp = case_inputs(name)
    print(f"no-alloc case {name}: n={len(pos)} m={len(smp)} cell={cs} range={rng_}", flush=True)
    sh = system.SpatialHashing(cs)
    ref = sh.find_samples(rng_, pos, smp)
    st, off, idx, total = noalloc(torch, sh, rng_, pos, smp)
    assert st == 0, N.lib().sph_last_error()
    assert total == len(ref.indices) and total > 0
    assert np.array_equal(off, ref.offsets)
    assert np.array_equal(idx, ref.indices)


def test_two_calls_in_a_row_second_smaller(torch_cuda):
    compare(torch_cuda, "random_large")
    compare(torch_cuda, "lattice_boundary")
    compare(torch_cuda, "wide_range_small")


def test_empty_inputs(torch_cuda):
    sh = system.SpatialHashing(0.1)
    print("no-alloc case empty inputs", flush=True)
    st, off, idx, total = noalloc(torch_cuda, sh, 0.1, np.zeros((0, 3)), np.zeros((5, 3)))
    assert st == 0 and total == 0 and np.array_equal(off, np.zeros(1, np.uint64))
    st, off, idx, total = noalloc(torch_cuda, sh, 0.1, np.zeros((4, 3)), np.zeros((0, 3)))
    assert st == 0 and total == 0 and np.array_equal(off, np.zeros(5, np.uint64))


def test_capacity_error_reports_total(torch_cuda):
    cs, rng_, pos, smp = case_inputs("wide_range_small")
    print("no-alloc case capacity", flush=True)
    sh = system.SpatialHashing(cs)
    ref = sh.find_samples(rng_, pos, smp)
    st, off, _, total = noalloc(torch_cuda, sh, rng_, pos, smp, cap=10)
    assert st == -4 and total == len(ref.indices)
    assert np.array_equal(off, ref.offsets)


def test_sizes_beyond_the_static_limit_are_refused(torch_cuda):
    sh = system.SpatialHashing(0.1)
    print("no-alloc case over the limit", flush=True)
    total = C.c_size_t(7)
    # refused before any device access: the pointers are never dereferenced
    buf = torch_cuda.zeros(8, dtype=torch_cuda.float64, device="cuda")
    torch_cuda.cuda.synchronize()
    st = N.lib().sph_neighbors_find_noalloc(sh._ctx.h, 0.1, 0.1, buf.data_ptr(), MAX_POINTS + 1,
                                            buf.data_ptr(), 3, buf.data_ptr(), buf.data_ptr(),
                                            8, C.byref(total))
    assert st == -5 and b"262144" in N.lib().sph_last_error()
