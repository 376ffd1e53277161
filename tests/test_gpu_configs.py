"""The measured configurations at their own sizes.

- The uneven batch at full size (24-replica 128^3, N = 160, FP32 grand-canonical string with
  lamellar -> gyroid-like ends, co-resident with 12 independent solves of m in {48^3, 64^3,
  96^3}, N in {64, 100, 128}): one round run concurrently (scft_iterate_concurrent) equals,
  bit for bit, the same round run context by context.
- Every distinct (m, N) context of that batch, FP32 and FP64, one props against the oracle at
  its full contour length, including the N = 128 solves whose f = 0.4 rounds to the odd block
  split n1 = 51, n2 = 77 (density.h:346, 405-422), and the 128^3 string.
- The string of the benchmark in FP64 (16 x 64^3, N = 128): props of three replicas against
  the oracle.
The oracle runs on the host's threads (16 at most), a few seconds per props at 96^3 and 128^3.
"""
import os
import sys

import numpy as np
import pytest

from bench import CFG, C4, initial_fields, uneven_contexts, uneven_specs
from oracle import orc
from scft_string_amd import Context, iterate_concurrent

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_uneven_batch_round_concurrent_equals_sequential():
    _say("uneven batch at full size: concurrent round")
    ctxs = uneven_contexts(True)
    try:
        iterate_concurrent(ctxs, [0] * len(ctxs))
        Wc = [c.get_fields() for c in ctxs]
    finally:
        for c in ctxs:
            c.close()
    _say("uneven batch at full size: sequential round")
    ctxs = uneven_contexts(True)
    try:
        for c in ctxs:
            c.string_iterate(0)
        Ws = [c.get_fields() for c in ctxs]
    finally:
        for c in ctxs:
            c.close()
    assert len(Wc) == 8
    for j, (a, b) in enumerate(zip(Wc, Ws)):
        assert np.array_equal(a, b), (j, float(np.max(np.abs(a - b))))


def _cases():
    out = []
    for kw, W in uneven_specs(True):
        label = f"{kw['m'][0]}cube_N{kw['N']}"
        out.append(pytest.param(kw, label, id=label))
    return out


@pytest.mark.parametrize("kw,label", _cases())
def test_uneven_batch_context_props_vs_oracle(kw, label):
    orc.set_threads(min(os.cpu_count() or 1, 16))
    W = [w for k, w in uneven_specs(True) if k["m"] == kw["m"] and k["N"] == kw["N"]][0]
    row = W.shape[0] // 2
    Wr = W[row:row + 1]
    m, D, N = kw["m"], kw["D"], kw["N"]
    n1 = int(round(N * C4["f"]))
    _say(f"{label}: props of row {row} (n1 = {n1}) against the oracle")
    ref = orc.Grid(m, D).props(Wr[0], N=N, alpha=C4["alpha"], phic=C4["phidb"], f=C4["f"], ens=C4["ens"])
    res = {}
    for prec in ("fp64", "fp32"):
        with Context(m, D, N, C4["f"], C4["alpha"], C4["ens"], C4["phidb"], C4["chi"], nrep=1,
                     precision=prec) as ctx:
            phi, lnQ, _ = ctx.props_batch(Wr)
        dq = abs(lnQ[0] - ref["lnQ"]) / abs(ref["lnQ"])
        dphi = float(np.max(np.abs(phi[0] - ref["phi"])) / np.max(np.abs(ref["phi"])))
        res[prec] = (dq, dphi)
    _say(f"{label}: FP64 rel dlnQ {res['fp64'][0]:.1e} dphi {res['fp64'][1]:.1e}; "
         f"FP32 rel dlnQ {res['fp32'][0]:.1e} dphi {res['fp32'][1]:.1e}")
    assert res["fp64"][0] <= 1e-10 and res["fp64"][1] <= 1e-10
    assert res["fp32"][0] <= 2e-5 and res["fp32"][1] <= 1e-4


def test_benchmark_string_fp64_props_vs_oracle():
    c = CFG
    R = c["nrep"]
    W = initial_fields(c["m"], c["D"], R)
    rows = [0, 8, 15]
    orc.set_threads(min(os.cpu_count() or 1, 16))
    g = orc.Grid(c["m"], c["D"])
    with Context(c["m"], c["D"], c["N"], c["f"], c["alpha"], c["ens"], c["phidb"], c["chi"], nrep=R,
                 precision="fp64", fix=c["fix"]) as ctx:
        phi, lnQ, phis = ctx.props_batch(W)
    for i in rows:
        ref = g.props(W[i], N=c["N"], alpha=c["alpha"], phic=c["phidb"], f=c["f"], ens=c["ens"])
        dq = abs(lnQ[i] - ref["lnQ"]) / max(1.0, abs(ref["lnQ"]))
        dphi = float(np.max(np.abs(phi[i] - ref["phi"])) / np.max(np.abs(ref["phi"])))
        dphis = float(np.max(np.abs(phis[i] - ref["phis"])))
        _say(f"16 x 64^3 N=128 FP64 row {i}: rel dlnQ {dq:.1e}, rel dphi {dphi:.1e}, dphis {dphis:.1e}")
        assert dq <= 1e-10 and dphi <= 1e-10 and dphis <= 1e-10
