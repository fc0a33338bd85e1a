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


def test_sequential_baseline_matches_batched():
    stream = serving.make_stream(n_requests=3, rate=1000.0, n_min=800, n_max=3000, steps=6,
                                 seed=4)
    seq = serving.SequentialServer().run(stream)
    bat = serving.BatchedServer(1 << 34, max_slots=8).run(stream)
    for rid in seq.results:
        assert seq.results[rid].digest == bat.results[rid].digest
        assert seq.results[rid].measurement == bat.results[rid].measurement
