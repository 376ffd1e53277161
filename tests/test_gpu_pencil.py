"""Line-pass (pencil) propagator vs the FP64 oracle, through the C-ABI.

The pencil kernels take any grid whose axes factor into 2, 3 and 5, including the
reference's shipped box m = 200 x 150 x 150, D = 16 x 12 x 12, N = 60, grand canonical
(input.dat:1-11). A (y, z) plane of that box does not fit in LDS. The FP64 bar is the same
1e-10 relative as for the plane kernels; FP32 reports its error.
"""
import sys

import numpy as np
import pytest

from oracle import orc
from scft_string_amd import Context

pytestmark = pytest.mark.gpu

TOL64 = 1e-10
PENCIL = {"family": "pencil"}  # these tests hold the pencil step; the split line step has its own


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


def _check64(m, D, N, f, alpha, ens, phidb, nrep, force):
    W = np.stack([_fields(m, D, seed=31 + i, mean_plus=0.3 * i) for i in range(nrep)])
    g = orc.Grid(m, D)
    with Context(m, D, N, f, alpha, ens, phidb, chi=15.0, nrep=nrep, precision="fp64", schedule=PENCIL) as ctx:
        assert ctx.fft_path == "pencil"
        phi, lnQ, _ = ctx.props_batch(W)
    for i in range(nrep):
        ref = g.props(W[i], N=N, alpha=alpha, phic=phidb, f=f, ens=ens)
        assert abs(lnQ[i] - ref["lnQ"]) <= TOL64 * max(1.0, abs(ref["lnQ"])), (i, lnQ[i], ref["lnQ"])
        for c in range(4):
            assert _rel(phi[i, c], ref["phi"][c]) <= TOL64, (i, c, _rel(phi[i, c], ref["phi"][c]))


CASES = [
    # m, D, N, f, alpha, ens, phidb, forced (grid also has plane/column kernels)
    ((16, 16, 16), (4.0, 4.0, 4.0), 40, 0.45, 0.1, 2, 0.9, True),
    ((12, 8, 16), (3.0, 2.5, 4.0), 30, 0.4, 0.2, 1, 0.7, True),
    ((20, 15, 15), (1.6, 1.2, 1.2), 24, 0.8, 0.1, 2, 0.3, False),
    ((15, 10, 9), (3.0, 2.0, 1.8), 21, 0.5, 0.1, 1, 0.8, False),
    ((1, 24, 30), (1.0, 3.0, 3.75), 20, 0.5, 0.1, 1, 0.8, False),
    ((40, 6, 50), (5.0, 1.0, 6.0), 17, 0.35, 0.15, 1, 0.6, False),
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: "x".join(map(str, c[0])) + f"_N{c[2]}_ens{c[5]}")
def test_pencil_props_fp64_matches_oracle(case, monkeypatch):
    m, D, N, f, alpha, ens, phidb, force = case
    _say("pencil props fp64", case)
    if force:
        monkeypatch.setenv("SCFT_PENCIL", "1")
    _check64(m, D, N, f, alpha, ens, phidb, 2, force)


def test_unit_axis_reduces_to_two_dimensions():
    """A field uniform along x gives the same lnQ on m0 = 1 as on m0 = 4 (the k_x = 0 plane)."""
    m2d, m3d, D = (1, 12, 10), (4, 12, 10), (2.0, 3.0, 2.5)
    W2 = _fields(m2d, D, seed=5)
    W3 = np.stack([np.tile(W2[c].reshape(1, -1), (4, 1)).reshape(-1) for c in range(2)])
    with Context(m2d, D, 20, 0.5, 0.1, 1, 0.8, chi=15.0, precision="fp64") as c2:
        _, lnQ2, _ = c2.props_batch(W2[None])
    with Context(m3d, D, 20, 0.5, 0.1, 1, 0.8, chi=15.0, precision="fp64") as c3:
        _, lnQ3, _ = c3.props_batch(W3[None])
    assert abs(lnQ2[0] - lnQ3[0]) <= 1e-13 * max(1.0, abs(lnQ3[0]))


@pytest.mark.parametrize("m", [(20, 15, 15), (40, 6, 50)])
def test_pencil_fp32_error_vs_oracle(m):
    D = tuple(0.2 * v for v in m)
    N, f = 30, 0.5
    W = _fields(m, D, seed=8)[None]
    ref = orc.Grid(m, D).props(W[0], N=N, alpha=0.1, phic=0.8, f=f, ens=1)
    with Context(m, D, N, f, 0.1, 1, 0.8, chi=15.0, precision="fp32", schedule=PENCIL) as ctx:
        assert ctx.fft_path == "pencil"
        phi, lnQ, _ = ctx.props_batch(W)
    dl = abs(lnQ[0] - ref["lnQ"])
    dphi = max(_rel(phi[0, c], ref["phi"][c]) for c in range(4))
    _say(f"pencil fp32 {m}: |dlnQ| {dl:.2e}, rel dphi {dphi:.2e}")
    assert dl < 1e-5 and dphi < 1e-4


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("fp32", 2e-5)])
def test_pencil_symmetric_operator_invariant(prec, tol):
    m, D = (20, 15, 15), (1.6, 1.2, 1.2)
    W = _fields(m, D, seed=9)[None]
    with Context(m, D, 40, 0.5, 0.1, 1, 0.8, chi=15.0, precision=prec) as ctx:
        ctx.props_batch(W)
        qq = ctx.qq_invariant(0)
    dev = float(np.max(np.abs(qq - qq[0])) / abs(qq[0]))
    _say(f"pencil {prec} invariant: max rel dev {dev:.2e}")
    assert dev <= tol


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("fp32", 1e-5)])
def test_pencil_and_plane_column_agree(prec, tol, monkeypatch):
    # 48 = 4 x 12 (z) and 6 x 8 (y) in the plane/column kernels (the default two-pass step), 8 x 2 x 3 in
    # the line kernels.
    m, D = (48, 48, 48), (5.0, 5.0, 5.0)
    W = np.stack([_fields(m, D, seed=40 + i) for i in range(3)])
    out = {}
    for pen in ("0", "1"):
        monkeypatch.setenv("SCFT_PENCIL", pen)
        with Context(m, D, 48, 0.45, 0.1, 1, 0.9, chi=18.0, nrep=3, precision=prec) as ctx:
            assert ctx.fft_path == ("pencil" if pen == "1" else "two-pass")
            out[pen] = ctx.props_batch(W)
    dl = float(np.max(np.abs(out["0"][1] - out["1"][1])))
    dphi = _rel(out["1"][0], out["0"][0])
    _say(f"pencil vs two-pass {prec}: |dlnQ| {dl:.2e}, rel dphi {dphi:.2e}")
    assert dl <= tol and dphi <= 10 * tol


@pytest.mark.parametrize("fam", ["pencil", "pencil-split"])
@pytest.mark.parametrize("m", [(128, 150, 160), (150, 12, 200), (96, 20, 96)])
@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("fp32", 1e-5)])
def test_compile_time_line_kernels_match_runtime(m, prec, tol, fam, monkeypatch):
    """Axes of lengths 96, 128, 150, 160, 192, 200, 256 run line kernels compiled for that length
    (fixed tiles, unrolled passes); SCFT_PENCIL_RT=1 forces the runtime-length plan. Both
    compute the same transform: props and the box gradient must agree."""
    D = tuple(0.08 * v for v in m)
    W = np.stack([_fields(m, D, seed=60 + i) for i in range(2)])
    out = {}
    for rt in ("0", "1"):
        monkeypatch.setenv("SCFT_PENCIL_RT", rt)
        _say(f"line kernels {m} {prec} runtime={rt}")
        with Context(m, D, 12, 0.45, 0.1, 1, 0.9, chi=18.0, nrep=2, precision=prec, schedule={"family": fam}) as ctx:
            assert ctx.fft_path == fam
            phi, lnQ, _ = ctx.props_batch(W)
            _, dq = ctx.dlnq_dbox(W)
            out[rt] = (phi, lnQ, dq)
    dl = float(np.max(np.abs(out["0"][1] - out["1"][1])))
    dphi = _rel(out["0"][0], out["1"][0])
    dd = _rel(out["0"][2], out["1"][2])
    _say(f"compile-time vs runtime {fam} {m} {prec}: |dlnQ| {dl:.2e}, rel dphi {dphi:.2e}, rel d(dlnQ/dD) {dd:.2e}")
    assert dl <= tol and dphi <= 10 * tol and dd <= 100 * tol


@pytest.mark.parametrize("fam", ["pencil", "default"])
def test_fp64_128_cube_vs_oracle(fam):
    """128^3 FP64: a plane does not fit in LDS, so this grid runs on the line kernels (by
    default the split line step)."""
    m, D, N = (128, 128, 128), (8.0, 8.0, 8.0), 16
    W = _fields(m, D, seed=12)[None]
    with Context(m, D, N, 0.45, 0.1, 1, 0.9, chi=18.0, precision="fp64", schedule={"family": fam}) as ctx:
        assert ctx.fft_path == ("pencil" if fam == "pencil" else "pencil-split")
        phi, lnQ, _ = ctx.props_batch(W)
    ref = orc.Grid(m, D).props(W[0], N=N, alpha=0.1, phic=0.9, f=0.45, ens=1)
    assert abs(lnQ[0] - ref["lnQ"]) <= TOL64 * max(1.0, abs(ref["lnQ"]))
    for c in range(4):
        assert _rel(phi[0, c], ref["phi"][c]) <= TOL64


def test_shipped_input_dat_box_fp64_vs_oracle():
    """The reference's own case: m = 200 150 150, D = 16 12 12, N = 60, ens = 2,
    phidb = 1.16e-5, f = 0.8, chi = 30 (input.dat:1-11), one replica, FP64, on the default
    (split) line step."""
    m, D, N = (200, 150, 150), (16.0, 12.0, 12.0), 60
    _say("input.dat box: GPU props")
    W = _fields(m, D, seed=2, amp=1.0)[None]
    with Context(m, D, N, 0.8, 0.1, 2, 1.16e-5, chi=30.0, precision="fp64") as ctx:
        assert ctx.fft_path == "pencil-split"
        phi, lnQ, _ = ctx.props_batch(W)
    _say("input.dat box: oracle props")
    orc.set_threads(16)
    ref = orc.Grid(m, D).props(W[0], N=N, alpha=0.1, phic=1.16e-5, f=0.8, ens=2)
    _say(f"input.dat box: lnQ {lnQ[0]!r} oracle {ref['lnQ']!r}")
    assert abs(lnQ[0] - ref["lnQ"]) <= TOL64 * max(1.0, abs(ref["lnQ"]))
    for c in range(4):
        assert _rel(phi[0, c], ref["phi"][c]) <= TOL64, (c, _rel(phi[0, c], ref["phi"][c]))
