"""The oracle against its committed golden trajectories
(tests/golden/oracle_trajectories.json, written by
tests/golden/make_oracle_trajectories.py): the scene generators reproduce the
recorded inputs digest, and the oracle reproduces every recorded state
digest, iteration count and error history. A change to either shows here
before it reaches a GPU comparison; tests/test_gpu_golden.py holds the HIP
path to the same records."""
import json
import os

import numpy as np
import pytest

import refcpu
from golden.make_oracle_trajectories import CASES, OUT, digest, inputs_digest, record

GOLD = json.load(open(OUT))["cases"]


@pytest.mark.parametrize("name", list(CASES))
def test_oracle_reproduces_golden_trajectory(name):
    make, steps = CASES[name]
    sc = make()
    gold = GOLD[name]
    assert inputs_digest(sc) == gold["inputs_sha256"]
    off, idx = refcpu.find_samples(sc.params.kernel_support_radius,
                                   sc.params.kernel_support_radius, sc.x, sc.x)
    assert digest(off) == gold["neighbors_step0"]["offsets"]
    assert digest(idx) == gold["neighbors_step0"]["indices"]
    o = refcpu.OracleSystem(sc)
    done = 0
    for s in steps:
        o.step(s - done)
        done = s
        got, want = record(o.state(), o.measure(), o.error_history()), gold["steps"][str(s)]
        for k in ("x", "v", "a", "V", "p", "solver_iterations"):
            assert got[k] == want[k], (s, k)
        np.testing.assert_allclose(got["density_error"], want["density_error"], rtol=1e-12)
        np.testing.assert_allclose(got["error_history"], want["error_history"], rtol=1e-12)
