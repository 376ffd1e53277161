"""tests/golden/c0_trajectory.npz is the oracle's own trajectory: its first iterations are
run again here on the oracle from the same seeded start and must agree to rounding (the GPU
test compares the full 200 iterations with the file)."""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))

from make_c0_trajectory import C0, oracle_state  # noqa: E402


def test_fixture_starts_like_the_oracle():
    ref = np.load(os.path.join(ROOT, "tests", "golden", "c0_trajectory.npz"))
    assert len(ref["errs"]) == C0["iters"] and len(ref["steps"]) == C0["iters"]
    st = oracle_state()
    for k in range(2):
        rep = st.iterate(k)
        assert abs(rep["errs"][0] - ref["errs"][k]) <= 1e-12 * ref["errs"][k]
        assert rep["steps"][0] == ref["steps"][k]
