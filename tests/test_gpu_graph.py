"""Graph replay of the contour loop (SCFT_GRAPH=1): each replica group's loop is captured
once per batch size on its own stream and replayed there. The replays run the same kernels
on the same data as direct launches, so props are bit-identical to a direct-launch context,
on every replay, across batch sizes (a new capture each) and with other work between."""
import sys

import numpy as np
import pytest

from scft_string_amd import Context

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _fields(m, D, R, seed=11):
    rng = np.random.default_rng(seed)
    M = int(np.prod(m))
    x = (np.arange(m[0]) * D[0] / m[0])[:, None, None] * np.ones(m)
    W = np.zeros((R, 2, M))
    for i in range(R):
        W[i, 0] = 1.5 * np.cos(2 * np.pi * x / D[0]).reshape(-1) + 0.3 * rng.standard_normal(M)
        W[i, 1] = 0.1 * rng.standard_normal(M)
    return W


@pytest.mark.parametrize("m,prec,streams", [((64, 64, 64), "fp32", 1), ((64, 64, 64), "fp32", 2),
                                            ((64, 64, 64), "fp32", 3), ((32, 32, 32), "fp64", 2)])
def test_graph_replay_matches_direct_launches(m, prec, streams, monkeypatch):
    D = (6.0, 6.0, 6.0)
    R = 5
    W = _fields(m, D, R)
    args = (m, D, 24, 0.45, 0.1, 1, 0.9, 18.0)
    sched = {"streams": streams}
    monkeypatch.setenv("SCFT_GRAPH", "0")
    with Context(*args, nrep=R, precision=prec, schedule=sched) as c0:
        ref5 = c0.props_batch(W)
        ref3 = c0.props_batch(W[:3])
    monkeypatch.setenv("SCFT_GRAPH", "1")
    with Context(*args, nrep=R, precision=prec, schedule=sched) as c1:
        assert "graph replay" in c1.engine_info, c1.engine_info
        runs = []
        for k in range(3):                      # capture, then two replays of the 5-row loop
            runs.append((5, c1.props_batch(W)))
            c1.set_fields(W)                    # other work on the context's stream between replays
            _ = c1.get_fields()
        runs.append((3, c1.props_batch(W[:3])))  # another batch size: a capture of its own
        runs.append((5, c1.props_batch(W)))      # the first capture again
        runs.append((3, c1.props_batch(W[:3])))
    for n, (phi, lnQ, phis) in runs:
        ref = ref5 if n == 5 else ref3
        assert np.array_equal(phi, ref[0]) and np.array_equal(lnQ, ref[1]) and np.array_equal(phis, ref[2]), n
    _say(f"graph replay {m} {prec} {streams} streams: {len(runs)} calls bit-identical to direct launches")
