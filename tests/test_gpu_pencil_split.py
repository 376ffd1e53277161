"""The split line step (scft_schedule family "pencil-split": line passes x, y, z, y per
contour step with the kinetic factor split by axis) against the FP64 oracle and the pencil
step, through the C-ABI. The same grids as the pencil tests: any axes factoring into 2, 3, 5,
compiled lengths (96 ... 256) and runtime-planned ones, including the reference's shipped box
m = 200 x 150 x 150 (input.dat:1-11)."""
import sys

import numpy as np
import pytest

from oracle import orc
from scft_string_amd import Context

pytestmark = pytest.mark.gpu

SPLIT = {"family": "pencil-split"}


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _fields(m, D, seed, amp=2.0, mean_plus=0.0):
    rng = np.random.default_rng(seed)
    M = int(np.prod(m))
    x = (np.arange(m[0]) * D[0] / m[0])[:, None, None] * np.ones(m)
    z = (np.arange(m[2]) * D[2] / m[2])[None, None, :] * np.ones(m)
    Wm = amp * (np.cos(2 * np.pi * x / D[0]) + 0.5 * np.sin(2 * np.pi * z / D[2])).reshape(-1)
    Wm += 0.3 * rng.standard_normal(M)
    Wp = mean_plus + 0.2 * rng.standard_normal(M)
    return np.stack([Wm, Wp])


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


CASES = [
    # m, D, N, f, alpha, ens, phidb
    ((16, 16, 16), (4.0, 4.0, 4.0), 40, 0.45, 0.1, 2, 0.9),
    ((12, 8, 16), (3.0, 2.5, 4.0), 30, 0.4, 0.2, 1, 0.7),
    ((20, 15, 15), (1.6, 1.2, 1.2), 24, 0.8, 0.1, 2, 0.3),
    ((15, 10, 9), (3.0, 2.0, 1.8), 21, 0.5, 0.1, 1, 0.8),
    ((1, 24, 30), (1.0, 3.0, 3.75), 20, 0.5, 0.1, 1, 0.8),
    ((40, 6, 50), (5.0, 1.0, 6.0), 17, 0.35, 0.15, 1, 0.6),
    ((96, 96, 96), (8.0, 8.0, 8.0), 6, 0.5, 0.125, 2, 0.7),   # compiled lengths
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: "x".join(map(str, c[0])) + f"_N{c[2]}_ens{c[5]}")
def test_pencil_split_props_fp64_matches_oracle(case):
    m, D, N, f, alpha, ens, phidb = case
    _say("pencil-split props fp64", case)
    W = np.stack([_fields(m, D, seed=31 + i, mean_plus=0.3 * i) for i in range(2)])
    g = orc.Grid(m, D)
    with Context(m, D, N, f, alpha, ens, phidb, chi=15.0, nrep=2, precision="fp64", schedule=SPLIT) as ctx:
        assert ctx.fft_path == "pencil-split"
        phi, lnQ, _ = ctx.props_batch(W)
    for i in range(2):
        ref = g.props(W[i], N=N, alpha=alpha, phic=phidb, f=f, ens=ens)
        dl = abs(lnQ[i] - ref["lnQ"]) / max(1.0, abs(ref["lnQ"]))
        dphi = max(_rel(phi[i, c], ref["phi"][c]) for c in range(4))
        _say(f"  row {i}: rel dlnQ {dl:.2e}, rel dphi {dphi:.2e}")
        assert dl <= 1e-10 and dphi <= 1e-10


def test_pencil_split_shipped_input_dat_box_vs_oracle():
    """The reference's own box and ensemble (m = 200 150 150, D = 16 12 12, ens = 2,
    phidb = 1.16e-5, f = 0.8, chi = 30), one replica, FP64, at N = 10 (the oracle's cost)."""
    m, D, N = (200, 150, 150), (16.0, 12.0, 12.0), 10
    W = _fields(m, D, seed=2, amp=1.0)[None]
    with Context(m, D, N, 0.8, 0.1, 2, 1.16e-5, chi=30.0, precision="fp64", schedule=SPLIT) as ctx:
        phi, lnQ, _ = ctx.props_batch(W)
    orc.set_threads(16)
    ref = orc.Grid(m, D).props(W[0], N=N, alpha=0.1, phic=1.16e-5, f=0.8, ens=2)
    dl = abs(lnQ[0] - ref["lnQ"]) / max(1.0, abs(ref["lnQ"]))
    dphi = max(_rel(phi[0, c], ref["phi"][c]) for c in range(4))
    _say(f"input.dat box, pencil-split: rel dlnQ {dl:.2e}, rel dphi {dphi:.2e}")
    assert dl <= 1e-10 and dphi <= 1e-10


@pytest.mark.parametrize("m,prec,tol", [((20, 15, 15), "fp64", 1e-12), ((20, 15, 15), "fp32", 2e-5),
                                        ((40, 30, 30), "fp32", 2e-5)])
def test_pencil_split_symmetric_operator_invariant(m, prec, tol):
    D = tuple(0.08 * v for v in m)
    W = _fields(m, D, seed=9)[None]
    with Context(m, D, 40, 0.5, 0.1, 1, 0.8, chi=15.0, precision=prec, schedule=SPLIT) as ctx:
        ctx.props_batch(W)
        qq = ctx.qq_invariant(0)
    dev = float(np.max(np.abs(qq - qq[0])) / abs(qq[0]))
    _say(f"pencil-split {m} {prec} invariant: max rel dev {dev:.2e}")
    assert dev <= tol


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("fp32", 2e-5)])
def test_pencil_split_agrees_with_pencil_and_checkpointed_store(prec, tol):
    """Against the pencil step on the same fields (3 replica groups on 2 streams), and the
    checkpointed q store (recompute windows resume through the x pass) against the full one."""
    m, D = (40, 30, 30), (4.0, 3.0, 3.0)
    W = np.stack([_fields(m, D, seed=50 + i) for i in range(3)])
    with Context(m, D, 30, 0.45, 0.1, 1, 0.8, chi=15.0, nrep=3, precision=prec, schedule={"family": "pencil"}) as c0:
        ref = c0.props_batch(W)
    out = {}
    for ck in (0, 6):
        with Context(m, D, 30, 0.45, 0.1, 1, 0.8, chi=15.0, nrep=3, precision=prec, ckpt=ck,
                     schedule={"family": "pencil-split", "streams": 2}) as c1:
            out[ck] = c1.props_batch(W)
    dl = float(np.max(np.abs(out[0][1] - ref[1])))
    dphi = _rel(out[0][0], ref[0])
    dck = _rel(out[6][0], out[0][0])
    _say(f"pencil-split vs pencil {prec}: |dlnQ| {dl:.2e}, rel dphi {dphi:.2e}; ckpt 6 vs full {dck:.2e}")
    assert dl <= tol * max(1.0, float(np.max(np.abs(ref[1])))) and dphi <= tol
    assert dck <= (1e-12 if prec == "fp64" else 1e-6)


@pytest.mark.parametrize("m", [(20, 15, 15), (40, 6, 50)])
def test_pencil_split_fp32_error_vs_oracle(m):
    D = tuple(0.2 * v for v in m)
    N, f = 30, 0.5
    W = _fields(m, D, seed=8)[None]
    ref = orc.Grid(m, D).props(W[0], N=N, alpha=0.1, phic=0.8, f=f, ens=1)
    with Context(m, D, N, f, 0.1, 1, 0.8, chi=15.0, precision="fp32", schedule=SPLIT) as ctx:
        phi, lnQ, _ = ctx.props_batch(W)
    dl = abs(lnQ[0] - ref["lnQ"])
    dphi = max(_rel(phi[0, c], ref["phi"][c]) for c in range(4))
    _say(f"pencil-split fp32 {m}: |dlnQ| {dl:.2e}, rel dphi {dphi:.2e}")
    assert dl < 1e-5 and dphi < 1e-4
