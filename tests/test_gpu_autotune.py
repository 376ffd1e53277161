"""The propagator autotuner (scft_params.autotune): at the first props it times every
kernel family x replica-group streams x graph replay on that batch and keeps the fastest.
Every candidate computes the same props (FP64: to rounding), so a tuned context must
agree with an untuned one. The choice must be the fastest candidate it logged.
"""
import re
import sys

import numpy as np
import pytest

from scft_string_amd import Context

pytestmark = [pytest.mark.gpu,
              pytest.mark.skip(reason="autotune is disabled: autotuned contexts crashed on the host (segmentation "
                                      "fault in the first props) and the cause is not found; scft_ctx_create "
                                      "refuses autotune = 1")]


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _fields(m, D, R, seed=2):
    rng = np.random.default_rng(seed)
    M = int(np.prod(m))
    x = (np.arange(m[0]) * D[0] / m[0])[:, None, None] * np.ones(m)
    W = np.zeros((R, 2, M))
    for i in range(R):
        W[i, 0] = 1.5 * np.cos(2 * np.pi * x / D[0]).reshape(-1) + 0.3 * rng.standard_normal(M)
    return W


@pytest.mark.parametrize("m,R", [((32, 32, 32), 3), ((16, 16, 16), 2)])
def test_autotuned_context_matches_default(m, R):
    D = (4.0, 4.0, 4.0)
    W = _fields(m, D, R)
    with Context(m, D, 30, 0.45, 0.1, 1, 0.8, 15.0, nrep=R, precision="fp64") as c0:
        phi0, lnQ0, _ = c0.props_batch(W)
    with Context(m, D, 30, 0.45, 0.1, 1, 0.8, 15.0, nrep=R, precision="fp64", autotune=1) as c1:
        phi1, lnQ1, _ = c1.props_batch(W)
        info = c1.engine_info
        phi2, lnQ2, _ = c1.props_batch(W)   # the chosen configuration, again
    _say(f"autotune {m}: {info}")
    assert "(autotuned)" in info
    assert np.max(np.abs(lnQ1 - lnQ0)) <= 1e-12 * np.max(np.abs(lnQ0))
    assert np.max(np.abs(phi1 - phi0)) <= 1e-12 * np.max(np.abs(phi0))
    assert np.array_equal(phi1, phi2)
    cands = re.findall(r"(\w[\w-]*)/(\d)/(graph|direct) ([0-9.]+);", info)
    assert len(cands) >= 4
    fam, ns, mode, _ = min(cands, key=lambda c: float(c[3]))
    assert info.startswith(f"{fam}, {ns} replica-group stream")
    assert ("graph replay" in info.split(";")[0]) == (mode == "graph")
