"""The HIP path against the committed golden trajectories
(tests/golden/oracle_trajectories.json): after each recorded step the fluid
state's binary64 bits hash to the oracle's recorded digests, with equal
iteration counts; the neighbor CSR of the initial positions
(sph_neighbors_find) hashes to the recorded one. No oracle runs here."""
import json

import numpy as np
import pytest

from golden.make_oracle_trajectories import CASES, OUT, digest

pytestmark = pytest.mark.gpu

GOLD = json.load(open(OUT))["cases"]


@pytest.fixture(scope="module")
def native():
    from sph_amd import system
    return system


@pytest.mark.parametrize("name", list(CASES))
def test_hip_path_reproduces_golden_trajectory(native, name):
    make, steps = CASES[name]
    sc = make()
    gold = GOLD[name]
    print(f"golden {name}: {sc.n} particles, steps {steps}", flush=True)
    h = sc.params.kernel_support_radius
    nl = native.SpatialHashing(h).find_samples(h, sc.x, sc.x)
    assert digest(nl.offsets) == gold["neighbors_step0"]["offsets"]
    assert digest(nl.indices) == gold["neighbors_step0"]["indices"]
    sim = native.System(sc)
    done = 0
    for s in steps:
        sim.step_forward_in_time(s - done)
        done = s
        st, m, want = sim.state(), sim.take_measurement(), gold["steps"][str(s)]
        for k in ("x", "v", "a", "V", "p"):
            assert digest(st[k]) == want[k], (s, k)
        assert int(m["solver_iterations"]) == want["solver_iterations"]
        np.testing.assert_allclose(m["density_error"], want["density_error"], rtol=1e-9)
