"""Plain SCFT of one replica at the reference's CPU-sized configuration, free-running on the
GPU for its full 200 iterations: m = 32^3, D = 4.0, N = 100, f = 0.5, chiN = 20,
phidb = 0.8, canonical, alpha = 0.1, dostring = 0 (solve_field, scft.cu:559-806, with one
replica and no string update). W- = 0.1 cos(2 pi 2x/D) + N(0, 0.01^2) from a seeded
generator, W+ = 0.

The oracle's run of the same 200 iterations is the fixture tests/golden/c0_trajectory.npz
(tests/golden/make_c0_trajectory.py; tests/test_c0_fixture.py re-runs its first iterations
on the oracle). Nothing is reset between iterations, so any difference that compounds shows:
the error norm of every iteration (relative 1e-9), the number of props every iteration runs
(equal), the fields at 256 points and their norms after the last iteration, and FreeE after
it (1e-9).
"""
import os
import sys

import numpy as np
import pytest

from scft_string_amd import Context

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))

from make_c0_trajectory import C0, c0_initial_fields  # noqa: E402


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_plain_scft_trajectory_matches_oracle_200_iterations():
    ref = np.load(os.path.join(ROOT, "tests", "golden", "c0_trajectory.npz"))
    c = C0
    M = int(np.prod(c["m"]))
    K = len(ref["errs"])
    assert K == c["iters"] == 200
    with Context(c["m"], c["D"], c["N"], c["f"], c["alpha"], c["ens"], c["phidb"], c["chi"], nrep=1,
                 precision="fp64", dostring=0) as ctx:
        ctx.set_fields(c0_initial_fields())
        for k in range(K):
            rg = ctx.string_iterate(k)
            eg, eo = float(rg["errs"][0]), float(ref["errs"][k])
            if k % 20 == 0 or k == K - 1:
                _say(f"k={k}: errs gpu {eg:.12e} oracle {eo:.12e}, props {int(rg['steps'][0])}/{int(ref['steps'][k])}")
            assert abs(eg - eo) <= 1e-9 * abs(eo), (k, eg, eo)
            assert int(rg["steps"][0]) == int(ref["steps"][k]), k
        W = ctx.get_fields()[0]
        fe = ctx.free_energy()
        Wfe = ctx.get_fields()[0]
    idx = ref["idx"]
    scale = max(np.max(np.abs(ref["Wm"])), np.max(np.abs(ref["Wp"])))
    dW = max(np.max(np.abs(W[:M][idx] - ref["Wm"])), np.max(np.abs(W[M:][idx] - ref["Wp"]))) / scale
    dWfe = max(np.max(np.abs(Wfe[:M][idx] - ref["Wm_fe"])), np.max(np.abs(Wfe[M:][idx] - ref["Wp_fe"]))) / scale
    dn = abs(np.linalg.norm(W[:M]) - ref["norm_Wm"]) / ref["norm_Wm"]
    _say(f"after {K} iterations: rel dW {dW:.2e} (after FreeE {dWfe:.2e}), norm {dn:.1e}, "
         f"FEs gpu {fe['FEs'][0]!r} oracle {float(ref['FEs'][0])!r}")
    assert dW <= 1e-9 and dWfe <= 1e-9 and dn <= 1e-9
    assert abs(fe["FEs"][0] - ref["FEs"][0]) <= 1e-9 * abs(ref["FEs"][0])
    assert abs(fe["phctot"][0] - ref["phctot"][0]) <= 1e-9
