"""The A homopolymer (scft_params.phihA) on the GPU, through the C-ABI.

FP64 props and d lnQ/dD against the oracle (1e-10). The A <-> B mirror symmetry
(negate W-, swap f <-> 1 - f) holds on the GPU. It maps the new A homopolymer onto the
reference's own B-homopolymer path. FP32 reports its error. Invalid settings are refused.
"""
import sys

import numpy as np
import pytest

from oracle import orc
from scft_string_amd import Context, ScftError

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _fields(m, D, R, seed):
    rng = np.random.default_rng(seed)
    M = int(np.prod(m))
    x = (np.arange(m[0]) * D[0] / m[0])[:, None, None] * np.ones(m)
    W = np.zeros((R, 2, M))
    for i in range(R):
        W[i, 0] = 1.5 * np.cos(2 * np.pi * x / D[0]).reshape(-1) + 0.3 * rng.standard_normal(M)
        W[i, 1] = 0.2 * rng.standard_normal(M) + 0.1 * i
    return W


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


CASES = [
    # m, D, N, f, ens, phidb, phihA
    ((16, 16, 16), (4.0, 4.0, 4.0), 20, 0.4, 1, 0.6, 0.25),
    ((16, 16, 16), (4.0, 4.0, 4.0), 20, 0.4, 2, 0.4, 0.7),
    ((20, 15, 15), (2.0, 1.5, 1.5), 30, 0.5, 1, 0.5, 0.3),   # line kernels, Nh = 3 (odd)
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: "x".join(map(str, c[0])) + f"_ens{c[4]}")
def test_props_and_gradient_fp64_match_oracle(case):
    m, D, N, f, ens, phidb, phihA = case
    _say("homopolymer A fp64", case)
    W = _fields(m, D, 2, seed=3)
    with Context(m, D, N, f, 0.1, ens, phidb, 15.0, nrep=2, precision="fp64", phihA=phihA) as ctx:
        phi, lnQ, phis = ctx.props_batch(W)
        _, d = ctx.dlnq_dbox(W)
    g = orc.Grid(m, D)
    g.set_phihA(phihA)
    for i in range(2):
        ref = g.props(W[i], N, 0.1, phidb, f, ens)
        assert abs(lnQ[i] - ref["lnQ"]) <= 1e-10 * max(1.0, abs(ref["lnQ"]))
        for c in range(4):
            assert _rel(phi[i, c], ref["phi"][c]) <= 1e-10, (i, c)
        assert abs(phis[i, 3] - ref["phis"][3]) <= 1e-10 * max(ref["phis"][3], 1e-3)
        _, rd = g.dlnq_dbox(W[i], N, 0.1, phidb, f, ens)
        assert _rel(d[i], rd) <= 1e-10, (i, d[i], rd)


@pytest.mark.parametrize("ens,phidb,pa,pb", [(1, 0.7, 0.3, 0.0), (2, 0.5, 1.0, 1.0)])
def test_mirror_symmetry_on_gpu(ens, phidb, pa, pb):
    m, D, N = (16, 16, 16), (4.0, 4.0, 4.0), 20
    W = _fields(m, D, 1, seed=9)
    Wm = W.copy()
    Wm[:, 0] *= -1
    with Context(m, D, N, 0.4, 0.1, ens, phidb, 15.0, precision="fp64", phihA=pa) as ca:
        phia, lqa, _ = ca.props_batch(W)
    with Context(m, D, N, 0.6, 0.1, ens, phidb, 15.0, precision="fp64", phihA=pb) as cb:
        phib, lqb, _ = cb.props_batch(Wm)
    assert abs(lqa[0] - lqb[0]) <= 1e-12 * abs(lqb[0])
    assert np.max(np.abs(phia[0, 0] - phib[0, 1])) <= 1e-12
    assert np.max(np.abs(phia[0, 1] - phib[0, 0])) <= 1e-12
    assert np.max(np.abs(phia[0, 2] - phib[0, 2])) <= 1e-12


def test_fp32_error():
    m, D, N, f, ens, phidb, phihA = (32, 32, 32), (4.0, 4.0, 4.0), 40, 0.4, 1, 0.6, 0.2
    W = _fields(m, D, 1, seed=5)
    with Context(m, D, N, f, 0.1, ens, phidb, 15.0, precision="fp32", phihA=phihA) as c32:
        phi, lnQ, _ = c32.props_batch(W)
    g = orc.Grid(m, D)
    g.set_phihA(phihA)
    ref = g.props(W[0], N, 0.1, phidb, f, ens)
    err = max(_rel(phi[0, c], ref["phi"][c]) for c in range(4))
    _say(f"homopolymer A fp32: |dlnQ| {abs(lnQ[0] - ref['lnQ']):.2e}, rel dphi {err:.2e}")
    assert abs(lnQ[0] - ref["lnQ"]) < 1e-5 and err < 1e-4


def test_invalid_settings_are_refused():
    m, D = (16, 16, 16), (4.0, 4.0, 4.0)
    with pytest.raises(ScftError):   # Nh = 5 > n1 = 4: q1 is not all-A up to Nh
        Context(m, D, 20, 0.2, 0.25, 1, 0.5, 15.0, phihA=0.1)
    with pytest.raises(ScftError):   # canonical fractions above 1
        Context(m, D, 20, 0.5, 0.1, 1, 0.8, 15.0, phihA=0.3)
    with pytest.raises(ScftError):
        Context(m, D, 20, 0.5, 0.1, 1, 0.5, 15.0, phihA=-0.1)
