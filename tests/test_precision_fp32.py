"""The precision choice, measured: binary32 arithmetic (with or without Kahan
summation) misses the 1e-5 relative bar on pressure, binary64 meets it.

SESPH's pressure is k * max(V0 / V - 1, 0) and the IISPH source term starts at
1 - V0 / V (sesph.rs:40-66, iisph.rs:175-250), with V = V0 / sum_j V0 W(r_ij)
(quantities.rs:30-58). At the fluid's working compression V0 / V - 1 is a few
1e-3, so the relative error of the sum is amplified by ~1 / (V0 / V - 1). The
test evaluates the cubic B-spline sum of interior particles of a compressed,
jittered lattice three ways: binary32 positions and arithmetic summed in
neighbor order, the same with Kahan compensation, and binary64; each against
an extended-precision (x87 80-bit) evaluation of the same sums.
"""
import math

import numpy as np
import pytest
from scipy.spatial import cKDTree


def lattice(n=16, dx=0.02, squeeze=0.997, jitter=1e-4, seed=7):
    rng = np.random.default_rng(seed)
    g = (np.arange(n) + 0.5) * dx * squeeze + 0.3
    x = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    return x + rng.uniform(-jitter, jitter, x.shape), dx


def w_sum(x, i, nb, h, dtype, kahan=False):
    """sum_j W(x_i - x_j) in `dtype` (W of kernel/mod.rs:22-36, powi as products)."""
    one, half, four = dtype(1), dtype(0.5), dtype(4)
    a = dtype(16.0 / math.pi) / (dtype(h) * dtype(h) * dtype(h))
    xi = x[i].astype(dtype)
    acc, c = dtype(0), dtype(0)
    for j in nb:
        r = xi - x[j].astype(dtype)
        q = np.sqrt((r[0] * r[0] + r[1] * r[1]) + r[2] * r[2]) / dtype(h)
        t1, t2 = max(one - q, dtype(0)), max(half - q, dtype(0))
        w = a * ((t1 * t1 * t1) - four * (t2 * t2 * t2))
        if kahan:
            y = w - c
            t = acc + y
            c = (t - acc) - y
            acc = t
        else:
            acc = acc + w
    return acc


@pytest.fixture(scope="module")
def sums():
    x, dx = lattice()
    h = 2 * dx
    tree = cKDTree(x)
    lo, hi = x.min(0) + 1.01 * h, x.max(0) - 1.01 * h
    inner = np.where(np.all((x > lo) & (x < hi), axis=1))[0]
    out = {}
    for name, dt, kahan in (("f32", np.float32, False), ("f32_kahan", np.float32, True),
                            ("f64", np.float64, False), ("ext", np.longdouble, False)):
        out[name] = np.array([w_sum(x, i, sorted(tree.query_ball_point(x[i], h)), h, dt, kahan)
                              for i in inner], dtype=np.longdouble)
    v0 = np.longdouble(dx) ** 3
    return {k: v0 * s - 1 for k, s in out.items()}   # V0 / V - 1 = V0 * sum W - 1


def rel_err(a, ref):
    m = ref > 0
    return np.abs(a[m] - ref[m]) / ref[m]


def test_the_fluid_is_compressed(sums):
    c = sums["ext"]
    assert np.all(c > 0) and 1e-3 < np.median(c) < 2e-2, np.median(c)


@pytest.mark.skipif(np.finfo(np.longdouble).eps >= np.finfo(np.float64).eps,
                    reason="no extended precision on this platform")
def test_binary32_misses_and_binary64_meets_the_pressure_bar(sums):
    ref = sums["ext"]
    e32, e32k, e64 = (rel_err(sums[k], ref) for k in ("f32", "f32_kahan", "f64"))
    print(f"relative error of V0/V - 1 (= p/k): f32 median {np.median(e32):.2e} max {e32.max():.2e}; "
          f"f32+Kahan median {np.median(e32k):.2e} max {e32k.max():.2e}; "
          f"f64 median {np.median(e64):.2e} max {e64.max():.2e}")
    assert np.median(e32) > 1e-5 and np.median(e32k) > 1e-5    # the bar is 1e-5 relative
    assert e64.max() < 1e-9
