"""Long-run parity: the HIP path against the oracle over whole runs.

The contract asks the average density error and the IISPH iteration counts
to match over long runs. Both sides evaluate every per-particle expression
in the reference's order in binary64, so the trajectories stay bit-identical
step after step; only cross-particle sums (the Jacobi error, the averages)
are grouped differently, and they agree to ~1e-16 relative. The checks:
every step's iteration counts (EQS1 and EQS2) and step size equal, every
step's measured and predicted average density error within 1e-9 relative,
and the full fluid state equal bit for bit at checkpoints.

Workloads: the SESPH dam break (8,000 particles, static samples) for all
1,000 steps of its configuration; a pool at rest (17,640 particles filling a
box wall to wall, 40 layers) for 300 steps under IISPH and under IISPHwOST,
where every solve carries the column's weight (13 and 17 Jacobi iterations
per step on average, up to 45); and a served stream of requests from 4,096
to 262,144 particles, each request's final state against the oracle run of
the same scene.
"""
import numpy as np
import pytest

import refcpu
from sph_amd import scenes, serving

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def native():
    from sph_amd import system
    return system


def close(a, b, rel=1e-9):
    return abs(a - b) <= rel * max(abs(a), abs(b), 1e-300)


def lockstep(native, sc, steps, checkpoints):
    o, g = refcpu.OracleSystem(sc), native.System(sc)
    its = []
    for k in range(1, steps + 1):
        o.step(1)
        g.step_forward_in_time(1)
        mo, mg = o.measure(), g.take_measurement()
        assert mg["solver_iterations"] == mo["solver_iterations"], (k, mg, mo)
        assert mg["eqs1_iterations"] == mo["eqs1_iterations"], (k, mg, mo)
        assert mg["time_step_size"] == mo["time_step_size"], k
        assert mg["time"] == mo["time"], k
        assert close(mg["density_error"], mo["density_error"]), (k, mg, mo)
        assert close(mg["predicted_density_error"], mo["predicted_density_error"]), (k, mg, mo)
        its.append(mo["solver_iterations"] + mo["eqs1_iterations"])
        if k in checkpoints:
            so, sg = o.state(), g.state()
            for f in ("x", "v", "a", "V", "p"):
                assert np.array_equal(sg[f], so[f]), (k, f)
            print(f"  step {k}: bit-identical; iterations so far mean {np.mean(its):.2f} "
                  f"max {max(its)}; density error {mo['density_error']:.4f} %", flush=True)
    return its


def test_dam_break_all_1000_steps(native):
    sc = scenes.dam_break_sesph()
    print(f"long run: SESPH dam break, {sc.n} particles, 1000 steps", flush=True)
    lockstep(native, sc, 1000, {100, 250, 500, 750, 1000})


@pytest.mark.parametrize("solver", ["IISPH", "IISPHwOST"])
def test_loaded_pool_300_steps(native, solver):
    sc = scenes.pool(solver=solver)
    print(f"long run: pool at rest, {solver}, {sc.n} particles, 300 steps", flush=True)
    its = lockstep(native, sc, 300, {50, 100, 200, 300})
    assert np.mean(its) > (4 if solver == "IISPH" else 8)  # loaded: above the minimum


def test_served_requests_match_the_oracle(native):
    """Requests of 4,096 to 262,144 particles (box and sphere containers, rest
    densities and viscosities across the stream's ranges) served together by
    continuous batching, 20 steps each; every finished request's state, step
    count, iteration count and density error equal the oracle's run of the
    same scene alone."""
    specs = [(4096, 800.0, 0.005, "box"), (12000, 1500.0, 0.05, "sphere"),
             (33000, 1000.0, 0.02, "box"), (90000, 1200.0, 0.01, "sphere"),
             (262144, 950.0, 0.03, "box")]
    stream = [serving.Request(k, 0.001 * k, scenes.serving_scene(n, rho, nu, shape, 100 + k), 20)
              for k, (n, rho, nu, shape) in enumerate(specs)]
    print(f"served stream: {[r.scene.n for r in stream]} particles, 20 steps each", flush=True)
    rep = serving.BatchedServer(8 << 30, max_slots=8, keep_state=True).run(stream)
    assert len(rep.results) == len(stream)
    for req in stream:
        got = rep.results[req.rid]
        o = refcpu.OracleSystem(req.scene)
        o.step(req.steps)
        so, mo = o.state(), o.measure()
        for f in ("x", "v", "a", "V", "p"):
            assert np.array_equal(got.state[f], so[f]), (req.rid, f)
        mg = got.measurement
        assert mg["steps"] == mo["steps"] == req.steps
        assert mg["solver_iterations"] == mo["solver_iterations"]
        assert close(mg["density_error"], mo["density_error"])
        print(f"  request {req.rid} ({req.scene.n} particles): bit-identical to the oracle",
              flush=True)
