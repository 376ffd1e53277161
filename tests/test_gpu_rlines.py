"""The split line step's register-FFT passes against the FP64 oracle and against the LDS
Stockham passes (SCFT_RLINES=0) on the same fields, through the C-ABI: the y and z passes
(k_ry, k_rz: the two-stage register FFT of the column pass along y and z, axis lengths 96,
128, 150, 160, 192, 256) and the x pass (the column kernel k_scol2 with the grid's (y, z)
extent at run time, x lengths 96 ... 256 and 200; FP64 without 256). Grids cover partial z tiles of the
y pass (m2 = 150), partial row tiles of the z pass (m0 m1 not a multiple of the tile), and
grids where only some axes have a register pass."""
import contextlib
import os
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


@contextlib.contextmanager
def _stockham_yz():
    old = os.environ.get("SCFT_RLINES")
    os.environ["SCFT_RLINES"] = "0"
    try:
        yield
    finally:
        if old is None:
            del os.environ["SCFT_RLINES"]
        else:
            os.environ["SCFT_RLINES"] = old


def _fields(m, D, seed, amp=2.0):
    rng = np.random.default_rng(seed)
    M = int(np.prod(m))
    x, y, z = np.meshgrid(*[np.arange(m[i]) * D[i] / m[i] for i in range(3)], indexing="ij")
    Wm = amp * (np.cos(2 * np.pi * x / D[0]) + 0.6 * np.sin(2 * np.pi * y / D[1])
                + 0.5 * np.cos(4 * np.pi * z / D[2])).reshape(-1)
    Wm += 0.3 * rng.standard_normal(M)
    Wp = 0.2 * rng.standard_normal(M)
    return np.stack([Wm, Wp])


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


CASES = [
    # m, D, N, f, ens: which passes are register ones
    ((8, 96, 150), (2.0, 8.0, 10.0), 12, 0.45, 2),    # y 96, z 150 (y tiles 32, 32, 32, 32, 22)
    ((6, 150, 96), (1.5, 10.0, 8.0), 12, 0.4, 1),     # y 150, z 96 (900 rows: last tile 4)
    ((5, 128, 100), (1.2, 8.0, 7.0), 10, 0.5, 1),     # y 128; z 100 on the Stockham pass
    ((6, 90, 160), (1.4, 7.0, 10.0), 10, 0.5, 2),     # y 90 on the Stockham pass; z 160 (540 rows)
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: "x".join(map(str, c[0])))
def test_register_line_passes_fp64_match_oracle(case):
    m, D, N, f, ens = case
    _say("register line passes fp64 vs oracle", case)
    W = np.stack([_fields(m, D, seed=61 + i) for i in range(2)])
    g = orc.Grid(m, D)
    with Context(m, D, N, f, 0.1, ens, 0.8, chi=15.0, nrep=2, precision="fp64", schedule=SPLIT) as ctx:
        assert ctx.fft_path == "pencil-split"
        phi, lnQ, _ = ctx.props_batch(W)
    for i in range(2):
        ref = g.props(W[i], N=N, alpha=0.1, phic=0.8, f=f, ens=ens)
        dl = abs(lnQ[i] - ref["lnQ"]) / max(1.0, abs(ref["lnQ"]))
        dphi = max(_rel(phi[i, c], ref["phi"][c]) for c in range(4))
        _say(f"  row {i}: rel dlnQ {dl:.2e}, rel dphi {dphi:.2e}")
        assert dl <= 1e-10 and dphi <= 1e-10


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("fp32", 3e-6)])
def test_register_line_passes_agree_with_stockham_passes(prec, tol):
    """The same props on the register and on the LDS Stockham y and z passes: 3 replicas in
    2 replica groups, the shipped box's y and z lengths (150) with a short x axis."""
    m, D, N = (12, 150, 150), (1.0, 12.0, 12.0), 16
    W = np.stack([_fields(m, D, seed=70 + i) for i in range(3)])
    sched = {"family": "pencil-split", "streams": 2}
    with Context(m, D, N, 0.8, 0.1, 2, 0.5, chi=20.0, nrep=3, precision=prec, schedule=sched) as c1:
        phi1, lnQ1, _ = c1.props_batch(W)
    with _stockham_yz():
        with Context(m, D, N, 0.8, 0.1, 2, 0.5, chi=20.0, nrep=3, precision=prec, schedule=sched) as c0:
            phi0, lnQ0, _ = c0.props_batch(W)
    dl = float(np.max(np.abs(lnQ1 - lnQ0)))
    dphi = _rel(phi1, phi0)
    _say(f"register vs Stockham y/z passes {prec}: |dlnQ| {dl:.2e}, rel dphi {dphi:.2e}")
    assert dl <= tol * max(1.0, float(np.max(np.abs(lnQ0)))) and dphi <= tol


@pytest.mark.parametrize("axis", [1, 2])
def test_register_line_passes_fp32_lamella(axis):
    """A lamella along y or z makes every line of that axis identical, the case where FP32
    roundings of the line sums add up coherently; the passes keep each line's sum in FP64."""
    m, D, N = (6, 150, 150), (0.6, 15.0, 15.0), 40
    M = int(np.prod(m))
    c = np.meshgrid(*[np.arange(m[i]) * D[i] / m[i] for i in range(3)], indexing="ij")
    other = 2 if axis == 1 else 1
    Wm = 1.5 * np.cos(2 * np.pi * c[other] / D[other]).reshape(-1)
    W = np.stack([Wm, np.zeros(M)])[None]
    out = {}
    for prec in ("fp64", "fp32"):
        with Context(m, D, N, 0.5, 0.1, 1, 0.8, chi=15.0, precision=prec, schedule=SPLIT) as ctx:
            out[prec] = ctx.props_batch(W)
    dl = abs(float(out["fp32"][1][0] - out["fp64"][1][0]))
    dphi = float(np.max(np.abs(out["fp32"][0] - out["fp64"][0])))
    _say(f"lamella varying along axis {other}: fp32 vs fp64 |dlnQ| {dl:.2e}, max |dphi| {dphi:.2e}")
    assert dl <= 3e-6 and dphi <= 1e-5


@pytest.mark.parametrize("m", [(200, 30, 24), (128, 20, 150), (200, 5, 5)], ids=lambda m: "x".join(map(str, m)))
def test_register_x_pass_fp32(m):
    """FP32: the x pass runs the two-pass step's column kernel with the grid's (y, z) extent
    at run time (x lengths 96 ... 256 and 200 = 20 x 10). Its FP32 props against an FP64
    context on the same fields, beside the FP32 error of the Stockham passes, and on a
    lamella along x (every x line identical, so FP32 line-sum roundings would add up
    coherently without the FP64 line sums)."""
    D = tuple(0.08 * v for v in m)
    N = 24
    W = np.stack([_fields(m, D, seed=80 + i) for i in range(2)])
    with Context(m, D, N, 0.45, 0.1, 1, 0.8, chi=15.0, nrep=2, precision="fp64", schedule=SPLIT) as c64:
        phi64, lnQ64, _ = c64.props_batch(W)
    with Context(m, D, N, 0.45, 0.1, 1, 0.8, chi=15.0, nrep=2, precision="fp32", schedule=SPLIT) as c1:
        phi1, lnQ1, _ = c1.props_batch(W)
    with _stockham_yz():
        with Context(m, D, N, 0.45, 0.1, 1, 0.8, chi=15.0, nrep=2, precision="fp32", schedule=SPLIT) as c0:
            phi0, lnQ0, _ = c0.props_batch(W)
    dl1, dp1 = float(np.max(np.abs(lnQ1 - lnQ64))), _rel(phi1, phi64)
    dl0, dp0 = float(np.max(np.abs(lnQ0 - lnQ64))), _rel(phi0, phi64)
    M = int(np.prod(m))
    x = np.meshgrid(*[np.arange(m[i]) * D[i] / m[i] for i in range(3)], indexing="ij")[0]
    Wl = np.stack([1.5 * np.cos(2 * np.pi * x / D[0]).reshape(-1), np.zeros(M)])[None]
    out = {}
    for prec in ("fp64", "fp32"):
        with Context(m, D, 40, 0.5, 0.1, 1, 0.8, chi=15.0, precision=prec, schedule=SPLIT) as ctx:
            out[prec] = ctx.props_batch(Wl)
    dll = abs(float(out["fp32"][1][0] - out["fp64"][1][0]))
    _say(f"x pass fp32 {m} vs fp64: register |dlnQ| {dl1:.2e} rel dphi {dp1:.2e}; Stockham |dlnQ| {dl0:.2e} "
         f"rel dphi {dp0:.2e}; x-lamella |dlnQ| {dll:.2e}")
    assert dl1 <= 3e-6 and dp1 <= 1e-5
    assert dll <= 3e-6


X_CASES = [
    # x lengths of the FP64 register x pass; (y, z) column counts that leave a partial
    # column tile
    ((150, 6, 5), (9.0, 1.0, 0.8), 12, 0.45, 2),
    ((128, 5, 9), (8.0, 0.8, 1.4), 10, 0.5, 1),
    ((160, 4, 6), (9.0, 0.7, 1.0), 10, 0.4, 1),
    ((192, 3, 5), (10.0, 0.5, 0.8), 8, 0.5, 2),
    ((200, 5, 5), (11.0, 0.8, 0.8), 10, 0.45, 1),    # 12-column tiles: 12, 12, 1
]


@pytest.mark.parametrize("case", X_CASES, ids=lambda c: "x".join(map(str, c[0])))
def test_register_x_pass_fp64_matches_oracle(case):
    m, D, N, f, ens = case
    _say("register x pass fp64 vs oracle", case)
    W = np.stack([_fields(m, D, seed=91 + i) for i in range(2)])
    g = orc.Grid(m, D)
    with Context(m, D, N, f, 0.1, ens, 0.8, chi=15.0, nrep=2, precision="fp64", schedule=SPLIT) as ctx:
        assert ctx.fft_path == "pencil-split"
        phi, lnQ, _ = ctx.props_batch(W)
    for i in range(2):
        ref = g.props(W[i], N=N, alpha=0.1, phic=0.8, f=f, ens=ens)
        dl = abs(lnQ[i] - ref["lnQ"]) / max(1.0, abs(ref["lnQ"]))
        dphi = max(_rel(phi[i, c], ref["phi"][c]) for c in range(4))
        _say(f"  row {i}: rel dlnQ {dl:.2e}, rel dphi {dphi:.2e}")
        assert dl <= 1e-10 and dphi <= 1e-10


def test_register_x_pass_fp64_checkpointed_store():
    """FP64, x = 150: the checkpointed q store (windows resume through the x pass) against
    the full store, and both against the Stockham passes."""
    m, D, N = (150, 10, 12), (9.0, 1.5, 1.8), 30
    W = np.stack([_fields(m, D, seed=97 + i) for i in range(3)])
    out = {}
    for ck in (0, 6):
        with Context(m, D, N, 0.45, 0.1, 1, 0.8, chi=15.0, nrep=3, precision="fp64", ckpt=ck,
                     schedule={"family": "pencil-split", "streams": 2}) as c1:
            out[ck] = c1.props_batch(W)
    with _stockham_yz():
        with Context(m, D, N, 0.45, 0.1, 1, 0.8, chi=15.0, nrep=3, precision="fp64", schedule=SPLIT) as c0:
            ref = c0.props_batch(W)
    dck = _rel(out[6][0], out[0][0])
    dl = float(np.max(np.abs(out[0][1] - ref[1])))
    dphi = _rel(out[0][0], ref[0])
    _say(f"x pass fp64 x=150: ckpt 6 vs full {dck:.2e}; vs Stockham |dlnQ| {dl:.2e} rel dphi {dphi:.2e}")
    assert dck <= 1e-12
    assert dl <= 1e-12 * max(1.0, float(np.max(np.abs(ref[1])))) and dphi <= 1e-12
