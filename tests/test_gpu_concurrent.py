"""Uneven solves on one GPU at once (scft_iterate_concurrent).

Contexts of different grids, contour lengths, precisions and modes iterate concurrently,
each from its own host thread and on its own streams. Every kernel and reduction is
deterministic, so the fields after concurrent iterations must equal, bit for bit, those
of the same iterations run one context after another.
"""
import sys

import numpy as np
import pytest

from scft_string_amd import Context, iterate_concurrent

pytestmark = pytest.mark.gpu

SETUPS = [
    # m, D, N, f, ens, phidb, chi, nrep, precision, dostring, fix
    ((16, 16, 16), (4.0, 4.0, 4.0), 20, 0.5, 1, 0.8, 15.0, 3, "fp64", 1, (1, 1)),   # string
    ((24, 24, 24), (4.0, 4.0, 4.0), 30, 0.45, 2, 0.5, 14.0, 2, "fp32", 0, (0, 0)),  # 2 independent solves
    ((20, 15, 15), (3.0, 2.2, 2.2), 24, 0.5, 1, 0.8, 15.0, 1, "fp64", 0, (0, 0)),   # line kernels
]


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _fields(m, D, R, seed):
    rng = np.random.default_rng(seed)
    M = int(np.prod(m))
    x = (np.arange(m[0]) * D[0] / m[0])[:, None, None] * np.ones(m)
    y = (np.arange(m[1]) * D[1] / m[1])[None, :, None] * np.ones(m)
    W = np.zeros((R, 2, M))
    for i in range(R):
        t = i / max(R - 1, 1)
        W[i, 0] = (2.0 * ((1 - t) * np.cos(2 * np.pi * x / D[0]) + t * np.cos(2 * np.pi * y / D[1]))).reshape(-1)
        W[i, 0] += 0.05 * rng.standard_normal(M)
    return W


def _make():
    ctxs = []
    for j, (m, D, N, f, ens, phidb, chi, R, prec, ds, fix) in enumerate(SETUPS):
        c = Context(m, D, N, f, 0.1, ens, phidb, chi, nrep=R, precision=prec, dostring=ds, fix=fix)
        c.set_fields(_fields(m, D, R, seed=100 + j))
        ctxs.append(c)
    return ctxs


def test_concurrent_iterations_equal_sequential():
    iters = 3
    _say("concurrent run")
    conc = _make()
    for k in range(iters):
        reps = iterate_concurrent(conc, [k] * len(conc))
    Wc = [c.get_fields() for c in conc]
    errs_c = [r["errs"] for r in reps]
    for c in conc:
        c.close()
    _say("sequential run")
    seq = _make()
    for c in seq:
        for k in range(iters):
            rep = c.string_iterate(k)
    Ws = [c.get_fields() for c in seq]
    for c in seq:
        c.close()
    for j in range(len(SETUPS)):
        assert np.array_equal(Wc[j], Ws[j]), (j, np.max(np.abs(Wc[j] - Ws[j])))
    assert np.array_equal(errs_c[-1], rep["errs"])


def test_concurrent_rejects_repeated_context():
    c = _make()[0]
    with pytest.raises(Exception):
        iterate_concurrent([c, c], [0, 0])
    c.close()
