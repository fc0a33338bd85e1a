"""Serving host logic on the CPU: the seeded request stream, the device-memory
estimate and slot/HBM-budget admission (no GPU calls)."""
import math

import numpy as np
import pytest

from sph_amd import scenes, serving


def test_stream_is_seeded_and_in_range():
    a = serving.make_stream(n_requests=12, rate=20.0, n_min=512, n_max=4096, steps=7, seed=3)
    b = serving.make_stream(n_requests=12, rate=20.0, n_min=512, n_max=4096, steps=7, seed=3)
    assert [r.arrival for r in a] == [r.arrival for r in b]
    for ra, rb in zip(a, b):
        assert np.array_equal(ra.scene.x, rb.scene.x)
        assert ra.scene.params == rb.scene.params
    arr = np.array([r.arrival for r in a])
    assert arr[0] == 0.0 and np.all(np.diff(arr) > 0)
    for r in a:
        p = r.scene.params
        assert r.steps == 7
        assert 0.4 * 512 <= r.scene.n <= 1.6 * 4096
        assert 800.0 <= r.scene.m[0] / p.rest_density_grid_spacing ** 3 <= 1500.0
        assert 0.005 <= p.fluid_viscosity <= 0.05 and p.boundary_viscosity == p.fluid_viscosity
        assert p.solver == "IISPH" and p.cfl and p.target_density_error == 0.1
    kinds = {r.scene.boundaries[0].kind for r in serving.make_stream(n_requests=16, n_min=512,
                                                                        n_max=1024, seed=1)}
    assert kinds == {"box", "sphere"}


def test_mixed_scheme_stream_changes_only_the_schemes():
    a = serving.make_stream(n_requests=9, n_min=1000, n_max=4000, steps=3, seed=5)
    b = serving.make_stream(n_requests=9, n_min=1000, n_max=4000, steps=3, seed=5,
                            schemes=serving.MIXED_SCHEMES)
    for k, (ra, rb) in enumerate(zip(a, b)):
        assert ra.arrival == rb.arrival and ra.steps == rb.steps
        assert np.array_equal(ra.scene.x, rb.scene.x)
        assert ra.scene.params.solver == "IISPH"
        solver, integrator = serving.MIXED_SCHEMES[k % 4]
        assert (rb.scene.params.solver, rb.scene.params.integrator) == (solver, integrator)
        # IISPHwOST pairs with TakePredicted only (setup/mod.rs:236-238)
        assert (solver == "IISPHwOST") == (integrator == "TakePredicted")


def test_stream_arrival_rate():
    # mean gap of the Poisson process is 1/rate (checked on a long seeded draw)
    rng = np.random.default_rng(0)
    gaps = rng.exponential(1.0 / 20.0, size=20000)
    assert abs(gaps.mean() - 0.05) < 0.002


@pytest.mark.parametrize("shape", ["box", "sphere"])
def test_serving_scene_fluid_inside_container(shape):
    sc = scenes.serving_scene(3000, 1000.0, 0.01, shape, 5)
    lat = sc.boundaries[0].sdf
    h = sc.params.kernel_support_radius
    if shape == "box":
        size = lat.x_max[0] - 3.1 * h
        lo, hi = np.zeros(3), np.full(3, size)
        d = np.minimum(sc.x - lo, hi - sc.x).min()
    else:
        r = lat.x_max[0] - 3.1 * h
        d = (r - np.linalg.norm(sc.x, axis=1)).min()
    assert d > sc.params.rest_density_grid_spacing  # at least one spacing from the wall


def test_containers_are_shared():
    a = scenes.serving_scene(3000, 900.0, 0.01, "box", 1)
    b = scenes.serving_scene(3100, 1200.0, 0.02, "box", 2)
    assert a.boundaries[0].sdf is b.boundaries[0].sdf


def test_device_estimate_grows_with_particles():
    small = scenes.serving_scene(2000, 1000.0, 0.01, "box", 1)
    big = scenes.serving_scene(16000, 1000.0, 0.01, "box", 1)
    es, eb = serving.estimate_device_bytes(small), serving.estimate_device_bytes(big)
    assert eb > es > small.n * serving.BYTES_PER_PARTICLE
    assert eb - es >= (big.n - small.n) * serving.BYTES_PER_PARTICLE


def test_native_admission_slots_budget_and_order():
    """The library's server admits in submission order under slots and the
    byte budget (no device is touched before a system is attached)."""
    from sph_amd import system
    srv = system.Server(budget_bytes=1000, max_slots=2)
    with pytest.raises(RuntimeError):
        srv.submit(1001, 5)            # never admissible
    with pytest.raises(RuntimeError):
        srv.submit(10, 0)              # no steps
    a, b, c = srv.submit(600, 5), srv.submit(500, 5), srv.submit(100, 5)
    assert srv.admit() == a
    assert srv.admit() is None         # b does not fit beside a; c waits behind b
    st = srv.stats()
    assert (st["queued"], st["admitted"], st["bytes"]) == (2, 1, 600)
    assert srv.poll() is None and srv.step()["world_steps"] == 0   # nothing resident
    srv.close()
    srv = system.Server(budget_bytes=1000, max_slots=2)
    t = [srv.submit(100, 1) for _ in range(3)]
    assert [srv.admit(), srv.admit(), srv.admit()] == [t[0], t[1], None]   # out of slots
    assert srv.stats()["admitted"] == 2
    srv.close()


def test_report_summary_percentiles():
    res = {k: serving.Result(k, 100 * (k + 1), 10, 0.1 * k, 0.1 * k, 0.1 * k + 1.0,
                             {"density_error": 0.01 * k, "predicted_density_error": 0.02 * k}, "")
           for k in range(4)}
    rep = serving.Report("batched", res, [1.0] * 99 + [5.0], 2.0, 1.0, 0.1, 4, 123)
    s = rep.summary()
    assert s["particle_steps"] == 10 * 100 * (1 + 2 + 3 + 4)
    assert math.isclose(s["throughput_particle_steps_per_s"], s["particle_steps"] / 2.0)
    assert s["p50_step_ms"] == 1.0 and 1.0 < s["p99_step_ms"] <= 5.0
    assert math.isclose(s["p50_request_s"], 1.0)
    assert math.isclose(s["density_error_pct_max"], 0.03)
    assert math.isclose(s["predicted_density_error_pct_max"], 0.06)


def test_arrival_bound_is_the_latest_arrival_plus_service():
    """The arrival bound: request r cannot finish before arrival_r plus its
    service time alone; the binding request is the one maximising that sum,
    and the ceiling is the stream's particle-steps over the bound."""
    from types import SimpleNamespace as NS
    reqs = [serving.Request(k, a, NS(n=n), s) for k, (a, n, s) in
            enumerate([(0.0, 100, 10), (1.0, 200, 5), (2.0, 50, 20)])]
    # sequential service times 3.0, 0.5, 0.25 s: r0 ends at 3.0, r1 at 3.5, r2 at 3.75
    times = [(0.0, 3.0), (3.0, 3.5), (3.5, 3.75)]
    res = {k: serving.Result(k, reqs[k].scene.n, reqs[k].steps, reqs[k].arrival, a, f, {}, "")
           for k, (a, f) in enumerate(times)}
    solo = serving.Report("sequential", res, [], 3.75, 3.75, 0.0, 1, 0)
    b = serving.arrival_bound(reqs, solo)
    # max(0 + 3.0, 1 + 0.5, 2 + 0.25) = 3.0, set by request 0
    assert b["makespan_lower_bound_s"] == 3.0 and b["binding_request"] == 0
    assert b["last_arrival_s"] == 2.0
    assert math.isclose(b["throughput_ceiling_particle_steps_per_s"], (1000 + 1000 + 1000) / 3.0)
    assert math.isclose(b["speedup_ceiling_vs_sequential"], 3.75 / 3.0)
    # a failed request does not count
    res[0].error = "solve did not converge"
    b = serving.arrival_bound(reqs, solo)
    assert b["makespan_lower_bound_s"] == 2.25 and b["binding_request"] == 2
