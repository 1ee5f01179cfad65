"""The oracle against its committed step fixtures (tests/golden/
oracle_steps_12x10x8.npz, oracle_c1_norms.json, written by
tests/golden/make_step_fixtures.py): every array bit for bit, the 64x64x32
norms to 1e-13 at steps 1 and 10.  The fixtures carry no reference output
(the reference has none for the step); they pin the oracle against later
changes and give the GPU tests outputs to meet without the oracle."""
import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
pytest.importorskip("pyoracle")
import make_step_fixtures as M  # noqa: E402

FIX = os.path.join(ROOT, "tests", "golden", "oracle_steps_12x10x8.npz")


@pytest.mark.parametrize("tag,bc", [("per", (0, 0, 1)), ("wall", (1, 1, 1))])
def test_small_boxes_reproduce_bitwise(tag, bc):
    want = np.load(FIX)
    got = M.small_case(bc)
    for k, v in got.items():
        w = want["%s_%s" % (tag, k)]
        assert w.shape == np.shape(v), k
        assert np.array_equal(w, v), k


def test_c1_norms_at_steps_1_and_10():
    want = json.load(open(os.path.join(ROOT, "tests", "golden", "oracle_c1_norms.json")))
    O = M.P.Oracle(64, 64, 32, bc=(0, 0, 1))
    st = O.init_state(M.P.make_spec([M.BUBBLE], noise_amp=1e-3, noise_seed=1))
    dt = O.compute_time_step(st, 0.8)
    assert dt == want["dt"]
    O.dim_switch = 1
    done = 0
    for target in (1, 10):
        for _ in range(target - done):
            O.time_step(st, dt)
        done = target
        I = O.interior(st)
        w = want["steps"][str(target)]
        for l in range(5):
            assert np.sqrt(np.mean(I[l] ** 2)) == pytest.approx(w["l2"][l], rel=1e-13)
            assert np.abs(I[l]).max() == pytest.approx(w["linf"][l], rel=1e-13)
