"""HIP path vs the FP64 oracle, through the C-ABI (libscft_hip.so).

FP64 bar: 1e-10 relative on Q (lnQ), the densities, the RMS exchange-chemical-potential
errors and the free energies. FP32 results carry their measured error against the oracle.
"""
import sys

import numpy as np
import pytest

from oracle import orc
from scft_string_amd import Context

pytestmark = pytest.mark.gpu

TOL64 = 1e-10


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _fields(m, D, seed, amp=2.0, mean_plus=0.0):
    rng = np.random.default_rng(seed)
    M = int(np.prod(m))
    x = (np.arange(m[0]) * D[0] / m[0])[:, None, None] * np.ones(m)
    Wm = amp * np.cos(2 * np.pi * x / D[0]).reshape(-1) + 0.3 * rng.standard_normal(M)
    Wp = mean_plus + 0.2 * rng.standard_normal(M)
    return np.stack([Wm, Wp])


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


CASES = [
    # m, D, N, f, alpha, ens, phidb
    ((8, 8, 8), (2.0, 2.0, 2.0), 20, 0.5, 0.1, 1, 0.8),
    ((12, 8, 16), (3.0, 2.5, 4.0), 30, 0.4, 0.2, 1, 0.7),
    ((16, 16, 16), (4.0, 4.0, 4.0), 40, 0.45, 0.1, 2, 0.9),
    ((8, 12, 10), (2.0, 3.0, 2.5), 25, 0.52, 0.12, 2, 0.3),
    ((32, 32, 32), (4.0, 4.0, 4.0), 50, 0.5, 0.1, 1, 0.8),
    ((24, 24, 24), (3.0, 3.0, 3.0), 33, 0.39, 0.1, 1, 0.6),
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: "x".join(map(str, c[0])) + f"_N{c[2]}_ens{c[5]}")
def test_props_fp64_matches_oracle(case):
    m, D, N, f, alpha, ens, phidb = case
    _say("props fp64", case)
    W = np.stack([_fields(m, D, seed=11 + i, mean_plus=0.4 * i) for i in range(3)])
    g = orc.Grid(m, D)
    with Context(m, D, N, f, alpha, ens, phidb, chi=15.0, nrep=3, precision="fp64") as ctx:
        phi, lnQ, phis = ctx.props_batch(W)
    for i in range(3):
        ref = g.props(W[i], N=N, alpha=alpha, phic=phidb, f=f, ens=ens)
        assert abs(lnQ[i] - ref["lnQ"]) <= TOL64 * max(1.0, abs(ref["lnQ"])), (i, lnQ[i], ref["lnQ"])
        for c in range(4):
            assert _rel(phi[i, c], ref["phi"][c]) <= TOL64, (i, c, _rel(phi[i, c], ref["phi"][c]))
        np.testing.assert_allclose(phis[i], ref["phis"], rtol=TOL64, atol=1e-14)


@pytest.mark.parametrize("m", [(16, 16, 16), (32, 32, 32)])
def test_props_fp32_error_vs_oracle(m):
    D = (4.0, 4.0, 4.0)
    N, f, alpha, ens, phidb = 40, 0.45, 0.1, 1, 0.85
    W = _fields(m, D, seed=3)[None]
    g = orc.Grid(m, D)
    ref = g.props(W[0], N=N, alpha=alpha, phic=phidb, f=f, ens=ens)
    with Context(m, D, N, f, alpha, ens, phidb, chi=15.0, nrep=1, precision="fp32") as ctx:
        phi, lnQ, phis = ctx.props_batch(W)
    dl = abs(lnQ[0] - ref["lnQ"])
    dphi = max(_rel(phi[0, c], ref["phi"][c]) for c in range(4))
    _say(f"fp32 {m}: |dlnQ|={dl:.3e} max rel dphi={dphi:.3e}")
    assert dl < 1e-5
    assert dphi < 1e-4


@pytest.mark.parametrize("fam", ["two-stage", "two-pass"])
@pytest.mark.parametrize("m,prec,tol", [
    ((32, 32, 32), "fp64", 1e-12), ((32, 32, 32), "fp32", 2e-5),
    ((48, 48, 48), "fp64", 1e-12),      # z split 4 x 12 (wave-local z exchanges)
    ((64, 64, 64), "fp64", 1e-12),      # plane_end compiled for two workgroups per CU
    ((64, 64, 64), "fp32", 2e-5),
    ((128, 128, 128), "fp32", 2e-5),    # z split 8 x 16, two stage-2 items per lane
], ids=lambda v: "x".join(map(str, v)) if isinstance(v, tuple) else str(v))
def test_symmetric_operator_invariant(m, prec, tol, fam):
    """<q1[s] q2[N-s]> is independent of s (fails for a wrong species switch, factor
    order or Richardson weight), on each plane-kernel geometry, for the four-launch
    two-stage step and the two-pass step."""
    D = tuple(v / 8.0 for v in m)
    N = 48
    W = _fields(m, D, seed=5, amp=3.0)[None]
    with Context(m, D, N, 0.4, 0.1, 1, 0.8, chi=15.0, nrep=1, precision=prec, schedule={"family": fam}) as ctx:
        assert ctx.fft_path == fam
        ctx.props_batch(W)
        qq = ctx.qq_invariant(0)
    spread = (qq.max() - qq.min()) / qq.mean()
    _say(f"{m} {prec} {fam}: qq spread {spread:.3e}")
    assert spread < tol


def test_solve_w_plus_matches_oracle():
    m, D = (12, 8, 16), (3.0, 2.5, 4.0)
    N, f, alpha, ens, phidb, chi = 30, 0.4, 0.1, 2, 0.5, 12.0
    R = 2
    W = np.stack([_fields(m, D, seed=21 + i) for i in range(R)])
    g = orc.Grid(m, D)
    M = g.M
    with Context(m, D, N, f, alpha, ens, phidb, chi=chi, nrep=R, precision="fp64") as ctx:
        ctx.set_fields(W)
        k = ctx.solve_w_plus(8, 1e-6)
        Wg = ctx.get_fields()
    for i in range(R):
        row = np.zeros(7 * M)
        row[:2 * M] = W[i].reshape(-1)
        dev = np.zeros(2 * M)
        kr = g.solve_field0(row, dev, chi, f, N, alpha, phidb, ens, 8, 1e-6)
        assert k[i] == kr
        assert _rel(Wg[i], row[:2 * M]) <= TOL64


def _string_setup(R, m, D, seed):
    M = int(np.prod(m))
    rng = np.random.default_rng(seed)
    x = (np.arange(m[0]) * D[0] / m[0])[:, None, None] * np.ones(m)
    base = np.cos(2 * np.pi * x / D[0]).reshape(-1)
    W = np.zeros((R, 2, M))
    for i in range(R):
        W[i, 0] = (0.6 + 0.5 * i) * base + 0.05 * rng.standard_normal(M)
        W[i, 1] = 0.05 * rng.standard_normal(M)
    return W


@pytest.mark.parametrize("springsteps", [0, 3])
def test_string_iteration_matches_oracle(springsteps):
    m, D = (8, 12, 10), (2.0, 3.0, 2.5)
    N, f, alpha, ens, phidb, chi = 24, 0.5, 0.125, 1, 0.8, 14.0
    R = 5
    W0 = _string_setup(R, m, D, seed=9)
    g = orc.Grid(m, D)
    run = dict(chi=chi, f=f, alpha=alpha, phidb=phidb, N=N, ens=ens, dostring=1, fix=(1, 1),
               springsteps=springsteps)
    st = orc.StringState(g, run, W0)
    with Context(m, D, N, f, alpha, ens, phidb, chi=chi, nrep=R, precision="fp64", fix=(1, 1),
                 springsteps=springsteps) as ctx:
        ctx.set_fields(W0)
        for k in range(2):
            _say("string iteration", k, "springsteps", springsteps)
            rep_g = ctx.string_iterate(k)
            rep_o = st.iterate(k)
            Wg = ctx.get_fields()
            assert _rel(Wg, st.W[:, :2 * g.M]) <= TOL64, _rel(Wg, st.W[:, :2 * g.M])
            np.testing.assert_array_equal(rep_g["steps"], rep_o["steps"])
            for key in ("errs", "errs3", "x", "dotWs", "dotWs2"):
                np.testing.assert_allclose(rep_g[key], rep_o[key], rtol=TOL64, atol=1e-13, err_msg=key)
            # teacher forcing: both continue from the GPU state
            st.W[:, :2 * g.M] = Wg


def test_free_energy_matches_oracle():
    m, D = (12, 8, 16), (3.0, 2.5, 4.0)
    N, f, alpha, ens, phidb, chi = 30, 0.5, 0.1, 1, 0.8, 14.0
    R = 4
    W0 = _string_setup(R, m, D, seed=4)
    g = orc.Grid(m, D)
    run = dict(chi=chi, f=f, alpha=alpha, phidb=phidb, N=N, ens=ens, dostring=1, fix=(1, 1))
    st = orc.StringState(g, run, W0)
    fe_o = st.free_energy()
    with Context(m, D, N, f, alpha, ens, phidb, chi=chi, nrep=R, precision="fp64", fix=(1, 1)) as ctx:
        ctx.set_fields(W0)
        fe_g = ctx.free_energy()
        Wg = ctx.get_fields()
    assert _rel(Wg, st.W[:, :2 * g.M]) <= TOL64
    for key in ("FEs", "FEs0", "FEs1", "phctot", "alf"):
        np.testing.assert_allclose(fe_g[key], fe_o[key], rtol=TOL64, atol=1e-12, err_msg=key)
    assert abs(fe_g["FE"] - fe_o["FE"]) <= TOL64 * abs(fe_o["FE"])


LARGE = [
    # m, D, N, f, alpha, ens, phidb, precision, tol (relative, vs the FP64 oracle)
    ((48, 48, 48), (5.0, 5.0, 5.0), 12, 0.5, 0.1, 1, 0.8, "fp64", TOL64),
    ((64, 64, 64), (6.0, 6.0, 6.0), 16, 0.45, 0.125, 1, 0.9, "fp64", TOL64),
    ((96, 96, 96), (8.0, 8.0, 8.0), 8, 0.5, 0.125, 2, 0.7, "fp64", TOL64),
    ((96, 96, 96), (8.0, 8.0, 8.0), 8, 0.5, 0.125, 1, 0.8, "fp32", 2e-5),
    ((128, 128, 128), (8.0, 8.0, 8.0), 8, 0.5, 0.125, 1, 0.8, "fp32", 2e-5),
]


@pytest.mark.parametrize("case", LARGE, ids=lambda c: f"{c[0][0]}cube_{c[7]}")
def test_props_large_grids_vs_oracle(case):
    m, D, N, f, alpha, ens, phidb, prec, tol = case
    _say("large props", m, prec)
    W = _fields(m, D, seed=17, amp=1.5)[None]
    ref = orc.Grid(m, D).props(W[0], N=N, alpha=alpha, phic=phidb, f=f, ens=ens)
    with Context(m, D, N, f, alpha, ens, phidb, chi=15.0, nrep=1, precision=prec) as ctx:
        phi, lnQ, phis = ctx.props_batch(W)
    dl = abs(lnQ[0] - ref["lnQ"]) / max(1.0, abs(ref["lnQ"]))
    dphi = max(_rel(phi[0, c], ref["phi"][c]) for c in range(4))
    _say(f"  {prec} {m}: rel dlnQ={dl:.2e} rel dphi={dphi:.2e}")
    assert dl <= tol and dphi <= tol


@pytest.mark.parametrize("pgrid", ["7", "256"])
def test_persistent_plane_mid_matches_per_plane(pgrid, monkeypatch):
    """96^2 FP32 planes run plane_mid as a persistent kernel: SCFT_PLANE_GRID workgroups (by
    default one per CU) each walk planes w, w + grid, ... and load the next plane during the
    current one's forward transform. SCFT_PLANE_GRID=0 runs one workgroup per plane. Both run
    the same transforms; FP32 rounding may differ where the compiler contracts differently.
    Three slots give 288 planes, so 7 workgroups take 41-42 planes each, and 256 take one or
    two."""
    m, D = (96, 96, 96), (6.0, 6.0, 6.0)
    W = np.stack([_fields(m, D, seed=70 + i, amp=1.5) for i in range(3)])
    out = {}
    for pg in ("0", pgrid):
        monkeypatch.setenv("SCFT_PLANE_GRID", pg)
        with Context(m, D, 24, 0.45, 0.1, 1, 0.9, chi=18.0, nrep=3, precision="fp32",
                     schedule={"family": "two-stage"}) as ctx:
            assert ctx.fft_path == "two-stage"
            assert ("persistent plane_mid" in ctx.engine_info) == (pg != "0"), ctx.engine_info
            phi, lnQ, _ = ctx.props_batch(W)
            qq = ctx.qq_invariant(2)
        out[pg] = (phi, lnQ, qq)
    dl = float(np.max(np.abs(out["0"][1] - out[pgrid][1])))
    dphi = _rel(out[pgrid][0], out["0"][0])
    qq = out[pgrid][2]
    dev = float(np.max(np.abs(qq - qq[0])) / abs(qq[0]))
    _say(f"persistent plane_mid (grid {pgrid}) vs per plane: |dlnQ| {dl:.2e}, rel dphi {dphi:.2e}, "
         f"invariant {dev:.2e}")
    assert dl <= 2e-5 and dphi <= 2e-5 and dev <= 1e-5


def test_96cube_fp64_stockham_and_line_kernels_agree(monkeypatch):
    """96^3 FP64 has no two-stage set, and its Stockham plane fills a CU's LDS, so it runs on
    the (split) line step by default; SCFT_LEGACY_FFT=1 keeps the Stockham kernels. Both
    agree."""
    m, D = (96, 96, 96), (8.0, 8.0, 8.0)
    W = _fields(m, D, seed=21, amp=1.5)[None]
    out = {}
    for leg in ("0", "1"):
        monkeypatch.setenv("SCFT_LEGACY_FFT", leg)
        with Context(m, D, 8, 0.5, 0.125, 2, 0.7, chi=15.0, precision="fp64") as ctx:
            assert ctx.fft_path == ("stockham" if leg == "1" else "pencil-split")
            out[leg] = ctx.props_batch(W)
    assert abs(out["0"][1][0] - out["1"][1][0]) <= 1e-12 * max(1.0, abs(out["1"][1][0]))
    assert _rel(out["0"][0], out["1"][0]) <= 1e-12


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_two_stage_and_stockham_paths_agree(prec, monkeypatch):
    """The two-stage register FFT and the LDS Stockham kernels compute the same props."""
    m, D = (32, 32, 32), (4.0, 4.0, 4.0)
    N, f, alpha, ens, phidb = 24, 0.5, 0.125, 1, 0.8
    W = _fields(m, D, seed=8)[None]
    out = {}
    for legacy in ("0", "1"):
        monkeypatch.setenv("SCFT_LEGACY_FFT", legacy)
        with Context(m, D, N, f, alpha, ens, phidb, chi=15.0, nrep=1, precision=prec) as ctx:
            out[legacy] = ctx.props_batch(W)
    tol = 1e-12 if prec == "fp64" else 1e-5
    assert _rel(out["0"][0], out["1"][0]) <= tol
    assert abs(out["0"][1][0] - out["1"][1][0]) <= tol * max(1.0, abs(out["1"][1][0]))


@pytest.mark.parametrize("fam,m,prec,width", [
    ("two-stage", (64, 64, 64), "fp32", 32), ("two-stage", (48, 48, 48), "fp64", 64),
    ("two-stage", (32, 32, 32), "fp64", 32), ("two-pass", (64, 64, 64), "fp32", 64),
    ("two-pass", (48, 48, 48), "fp64", 64), ("two-pass", (32, 32, 32), "fp64", 64)])
def test_column_tile_widths_agree(m, prec, width, fam, monkeypatch):
    """The column pass at either tile width (the autotuner may pick either; `width` is the
    family's non-default one on that grid) transforms every column with the same arithmetic:
    props are bit-identical, for the two-stage and the two-pass step."""
    D = (5.0, 5.0, 5.0)
    W = np.stack([_fields(m, D, seed=30 + i) for i in range(2)])
    out = {}
    for ct in ("0", str(width)):
        monkeypatch.setenv("SCFT_COLTILE", ct)
        with Context(m, D, 24, 0.45, 0.1, 1, 0.9, chi=15.0, nrep=2, precision=prec, schedule={"family": fam}) as ctx:
            assert ctx.fft_path == (fam if ct == "0" else f"{fam}-c{width}")
            out[ct] = ctx.props_batch(W)
    assert np.array_equal(out["0"][0], out[str(width)][0]) and np.array_equal(out["0"][1], out[str(width)][1])


@pytest.mark.parametrize("m,prec,nt", [((64, 64, 64), "fp32", 256), ((64, 64, 64), "fp64", 256),
                                       ((96, 96, 96), "fp32", 384)])
def test_plane_workgroup_sizes_agree(m, prec, nt, monkeypatch):
    """The two-stage plane kernels at half the default workgroup size (SCFT_PLANE_NT; each
    thread then takes two items of every stage) run the same transforms on the same data:
    props agree to rounding (the compiler may contract differently per instantiation), and
    the symmetric-operator invariant holds."""
    D = (6.0, 6.0, 6.0)
    W = np.stack([_fields(m, D, seed=40 + i, amp=1.5) for i in range(2)])
    out = {}
    for pn in ("0", str(nt)):
        monkeypatch.setenv("SCFT_PLANE_NT", pn)
        with Context(m, D, 24, 0.45, 0.1, 1, 0.9, chi=18.0, nrep=2, precision=prec,
                     schedule={"family": "two-stage"}) as ctx:
            assert ("%d-thread plane workgroups" % nt in ctx.engine_info) == (pn != "0"), ctx.engine_info
            phi, lnQ, _ = ctx.props_batch(W)
            qq = ctx.qq_invariant(1)
        out[pn] = (phi, lnQ, qq)
    tol = 1e-12 if prec == "fp64" else 2e-5
    dl = float(np.max(np.abs(out["0"][1] - out[str(nt)][1])))
    dphi = _rel(out[str(nt)][0], out["0"][0])
    qq = out[str(nt)][2]
    dev = float(np.max(np.abs(qq - qq[0])) / abs(qq[0]))
    _say(f"{m} {prec} plane workgroups {nt} vs default: |dlnQ| {dl:.2e}, rel dphi {dphi:.2e}, invariant {dev:.2e}")
    assert dl <= tol and dphi <= tol and dev <= (1e-12 if prec == "fp64" else 1e-5)
