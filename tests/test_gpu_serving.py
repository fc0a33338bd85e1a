"""Continuous-batching service on the GPU: every request's result is the same
alone and in the mix (bit-identical), under slot and HBM-budget admission."""
import numpy as np
import pytest

from sph_amd import serving

pytestmark = pytest.mark.gpu


def solo_result(req):
    from sph_amd import system
    sim = system.System(req.scene)
    sim.step_forward_in_time(req.steps)
    return sim.state(), sim.take_measurement()


def test_batched_stream_matches_solo_runs():
    stream = serving.make_stream(n_requests=7, rate=400.0, n_min=800, n_max=6000, steps=12,
                                 seed=11)
    est = sorted(serving.estimate_device_bytes(r.scene) for r in stream)
    budget = est[-1] + est[-2] + est[-3]    # about three scenes resident at a time
    srv = serving.BatchedServer(budget, max_slots=4, keep_state=True)
    rep = srv.run(stream)
    assert set(rep.results) == {r.rid for r in stream}
    assert 2 <= rep.peak_resident <= 4 and rep.peak_bytes <= budget
    for req in stream:
        st, meas = solo_result(req)
        got = rep.results[req.rid]
        assert got.steps == req.steps and got.n == len(st["x"])
        for k in ("x", "v", "a", "V", "p"):
            assert np.array_equal(got.state[k], st[k]), (req.rid, k)
        assert got.measurement == meas
        assert got.measurement["density_error"] <= 0.1 + 1e-12
    s = rep.summary()
    assert s["requests"] == 7 and s["p99_step_ms"] >= s["p50_step_ms"] > 0


def test_mixed_solvers_stream_matches_solo_runs():
    """Requests of the four pressure solvers (IISPH, IISPHwOST with
    TakePredicted, SESPH with ExplicitEuler, SESPHwSplitting) in one stream:
    the server keeps one world per solver and steps them all in each server
    step; every request equals its solo run bit for bit."""
    from sph_amd import scenes
    kinds = [("IISPH", "EulerCromer"), ("IISPHwOST", "TakePredicted"),
             ("SESPH", "ExplicitEuler"), ("SESPHwSplitting", "EulerCromer"),
             ("IISPH", "EulerCromer"), ("SESPHwSplitting", "EulerCromer")]
    stream = []
    for k, (solver, integ) in enumerate(kinds):
        sc = scenes.serving_scene(1500 + 700 * k, 1000.0, 0.01, "box" if k % 2 else "sphere", 40 + k)
        sc.params.solver, sc.params.integrator = solver, integ
        stream.append(serving.Request(k, 0.0, sc, 8 + k))
    print(f"mixed-solver stream: {[r.scene.n for r in stream]} particles", flush=True)
    rep = serving.BatchedServer(1 << 34, max_slots=8, keep_state=True).run(stream)
    assert set(rep.results) == {r.rid for r in stream}
    assert rep.peak_resident == len(stream)  # all admitted at once, four worlds
    for req in stream:
        st, meas = solo_result(req)
        got = rep.results[req.rid]
        for k in ("x", "v", "a", "V", "p"):
            assert np.array_equal(got.state[k], st[k]), (req.rid, k)
        assert got.measurement == meas, req.rid


def test_sequential_baseline_matches_batched():
    stream = serving.make_stream(n_requests=3, rate=1000.0, n_min=800, n_max=3000, steps=6,
                                 seed=4)
    seq = serving.SequentialServer().run(stream)
    bat = serving.BatchedServer(1 << 34, max_slots=8).run(stream)
    for rid in seq.results:
        assert seq.results[rid].digest == bat.results[rid].digest
        assert seq.results[rid].measurement == bat.results[rid].measurement


TWINS = np.array([3, 77, 200, 200, 411])


def twin_scene():
    """Particles at exactly the same positions never separate under IISPH,
    so every pressure solve stalls above its target."""
    from sph_amd import scenes
    sc = scenes.small_iisph_box(side=8, jitter=2e-3)
    for k in ("x", "v", "m", "fluid_id"):
        setattr(sc, k, np.concatenate([getattr(sc, k), getattr(sc, k)[TWINS]]))
    return sc


def test_world_member_failure_is_confined():
    """A world member whose solve exceeds its iteration cap fails that world
    step alone: the call reports it (sph_system_status names it), the other
    member has completed the step and equals its solo run bit for bit, and
    the world steps on after the failed member is detached."""
    from sph_amd import _native as N
    from sph_amd import scenes, system
    pool = scenes.pool(solver="IISPH")  # 3 iterations at first, more once loaded
    other = scenes.small_iisph_box(side=10, jitter=2e-3)
    a = system.System(pool, init=False)
    a.set_max_iterations(3)
    a.init()
    b = system.System(other)
    w = system.World([a, b])
    steps = 0
    print("world with a capped member: stepping until it fails", flush=True)
    with pytest.raises(N.SphError, match="did not converge within 3 iterations"):
        while steps < 400:
            w.step(1)
            steps += 1
    steps += 1  # the failing step completed for the other member
    assert a.status()[0] == -5 and b.status()[0] == 0
    print(f"  member failed in step {steps}", flush=True)
    w.detach(a)
    w.step(5)
    w.detach(b)
    solo = system.System(other)
    solo.step_forward_in_time(steps + 5)
    sb, ss = b.state(), solo.state()
    for k in ("x", "v", "a", "V", "p"):
        assert np.array_equal(sb[k], ss[k]), k
    w.close()


def test_served_stream_isolates_failing_requests():
    """One request fails in System::new (twin particles, cap 50), one fails
    mid-run (a pool capped at 3 iterations once it is loaded); both are
    reported failed with their status and the other requests finish
    bit-identical to their solo runs, in the batched and in the sequential
    server."""
    from sph_amd import scenes
    good = serving.make_stream(n_requests=2, rate=1000.0, n_min=1500, n_max=4000, steps=30,
                               seed=21)
    bad = [serving.Request(10, 0.0, twin_scene(), 30, max_iterations=50),
           serving.Request(11, 0.0, scenes.pool(solver="IISPH"), 150, max_iterations=3)]
    stream = good + bad
    print(f"stream with failing requests: {[r.scene.n for r in stream]}", flush=True)
    for server in (serving.BatchedServer(1 << 34, max_slots=8, keep_state=True),
                   serving.SequentialServer(keep_state=True)):
        rep = server.run(stream)
        assert set(rep.results) == {r.rid for r in stream}
        assert set(rep.failed) == {10, 11}
        for rid in (10, 11):
            assert "did not converge" in rep.results[rid].error
        for req in good:
            st, meas = solo_result(req)
            got = rep.results[req.rid]
            assert got.error is None
            for k in ("x", "v", "a", "V", "p"):
                assert np.array_equal(got.state[k], st[k]), (req.rid, k)
        assert rep.summary()["failed_requests"] == 2


def test_device_bytes_held_and_world_scratch_released():
    """The library counts the device bytes its buffers hold: a system's
    creation raises the count and closing it returns it. A world's launch
    buffers, which only grow, are freed once a large request has left and
    the residents need less than a quarter of them; the estimates admission
    uses stay above what a resident request holds."""
    import gc
    from sph_amd import scenes, system
    gc.collect()                     # systems of earlier tests free their buffers first
    b0 = system.device_bytes_held()
    small = scenes.serving_scene(4096, 1000.0, 0.01, "box", 1)
    sim = system.System(small)
    assert system.device_bytes_held() > b0
    sim._ctx.close()
    assert system.device_bytes_held() == b0
    big = scenes.serving_scene(300000, 1000.0, 0.01, "box", 2)
    srv = system.Server(1 << 36, 4)
    sims = {}
    for sc, steps in ((big, 2), (small, 8)):
        t = srv.submit(serving.estimate_device_bytes(sc), steps)
        assert srv.admit() == t
        sims[t] = system.System(sc)
        srv.attach(t, sims[t])
    peak = srv.step()["device_bytes"]
    srv.step()                       # the large request finishes and leaves
    got = srv.poll()
    assert got is not None and len(got[1]) == big.n
    got[1]._ctx.close()
    after = srv.step()["device_bytes"]
    held_small = after - b0
    assert after < 0.5 * peak, (after, peak)
    assert held_small <= serving.estimate_device_bytes(small) + (8 << 20), held_small
    srv.close()
    for s in sims.values():
        if s._ctx.h:
            s._ctx.close()


def test_server_options_before_the_first_request():
    """A server's world options can be set before any request is attached
    (the world is created on first use); with graph replay off the stream's
    results are those of the default server, bit for bit."""
    stream = serving.make_stream(n_requests=3, rate=1000.0, n_min=800, n_max=3000, steps=6,
                                 seed=5)
    a = serving.BatchedServer(1 << 34, max_slots=8, keep_state=True).run(stream)
    b = serving.BatchedServer(1 << 34, max_slots=8, keep_state=True, graphs=False).run(stream)
    for req in stream:
        ra, rb = a.results[req.rid], b.results[req.rid]
        assert ra.digest == rb.digest and ra.measurement == rb.measurement, req.rid
