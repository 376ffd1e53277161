"""Box derivative d lnQ / dD (the adjoint of props) on the GPU, through the C-ABI.

FP64 against the oracle's orc_dlnq_dbox (itself pinned by central differences of
orc_props): 1e-10 relative to the largest component. The GPU result is also checked
against 4-point central differences of the GPU's own lnQ, with the box moved by
scft_set_box. FP32 reports its error.
"""
import sys

import numpy as np
import pytest

from oracle import orc
from scft_string_amd import Context

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _fields(m, D, seed):
    rng = np.random.default_rng(seed)
    M = int(np.prod(m))
    x = (np.arange(m[0]) * D[0] / m[0])[:, None, None] * np.ones(m)
    y = (np.arange(m[1]) * D[1] / m[1])[None, :, None] * np.ones(m)
    z = (np.arange(m[2]) * D[2] / m[2])[None, None, :] * np.ones(m)
    Wm = (2.0 * np.cos(2 * np.pi * x / D[0]) + np.cos(2 * np.pi * y / D[1])
          + 0.5 * np.sin(2 * np.pi * z / D[2])).reshape(-1)
    Wm += 0.3 * rng.standard_normal(M)
    Wp = 0.2 * rng.standard_normal(M)
    return np.stack([Wm, Wp])


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


CASES = [
    # m, D, N, f, alpha, ens, phidb
    ((16, 16, 16), (4.0, 3.6, 4.4), 40, 0.45, 0.1, 1, 0.9),      # plane kernels for props
    ((12, 8, 16), (3.1, 2.2, 3.7), 21, 0.4, 0.1, 2, 0.3),
    ((20, 15, 15), (2.0, 1.5, 1.6), 24, 0.55, 0.2, 1, 0.6),      # line kernels, odd n1
    ((32, 1, 1), (4.0, 1.0, 1.0), 30, 0.5, 0.1, 1, 0.8),        # 1-D box
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: "x".join(map(str, c[0])) + f"_ens{c[5]}")
def test_dlnq_dbox_fp64_matches_oracle(case):
    m, D, N, f, alpha, ens, phidb = case
    _say("dlnq_dbox fp64", case)
    W = np.stack([_fields(m, D, seed=20 + i) for i in range(3)])
    with Context(m, D, N, f, alpha, ens, phidb, chi=15.0, nrep=3, precision="fp64") as ctx:
        lnQ, d = ctx.dlnq_dbox(W)
    g = orc.Grid(m, D)
    for i in range(3):
        rl, rd = g.dlnq_dbox(W[i], N, alpha, phidb, f, ens)
        assert abs(lnQ[i] - rl) <= 1e-10 * max(1.0, abs(rl))
        assert _rel(d[i], rd) <= 1e-10, (i, d[i], rd)


def test_dlnq_dbox_matches_gpu_central_differences():
    m, D, N, f, alpha, ens, phidb = (12, 8, 16), (3.1, 2.2, 3.7), 21, 0.4, 0.1, 1, 0.7
    W = _fields(m, D, seed=4)[None]
    with Context(m, D, N, f, alpha, ens, phidb, chi=15.0, precision="fp64") as ctx:
        _, d = ctx.dlnq_dbox(W)
        fd = np.zeros(3)
        for i in range(3):
            h = 1e-3 * D[i]
            v = []
            for k in (2, 1, -1, -2):
                Dk = list(D)
                Dk[i] += k * h
                ctx.set_box(Dk)
                v.append(ctx.props_batch(W)[1][0])
            fd[i] = (-v[0] + 8 * v[1] - 8 * v[2] + v[3]) / (12 * h)
        ctx.set_box(D)
        _, d2 = ctx.dlnq_dbox(W)
    _say(f"gpu adjoint {d[0]} fd {fd}")
    assert _rel(d[0], fd) <= 1e-8
    assert np.array_equal(d, d2)


@pytest.mark.parametrize("m", [(16, 16, 16), (20, 15, 15), (64, 64, 64)])
def test_dlnq_dbox_fp32_error(m):
    D = tuple(0.25 * v for v in m)
    N, f, alpha, ens, phidb = 32, 0.5, 0.1, 1, 0.8
    W = _fields(m, D, seed=6)[None]
    with Context(m, D, N, f, alpha, ens, phidb, chi=15.0, precision="fp64") as c64:
        _, d64 = c64.dlnq_dbox(W)
    with Context(m, D, N, f, alpha, ens, phidb, chi=15.0, precision="fp32") as c32:
        _, d32 = c32.dlnq_dbox(W)
    err = _rel(d32, d64)
    _say(f"dlnq_dbox fp32 {m}: rel error {err:.2e}  ({d32[0]} vs {d64[0]})")
    assert err <= 1e-4


# ---------------------------------------------------------------- saddle point, relaxation
LAM = dict(m=(32, 1, 1), N=40, f=0.5, alpha=0.1, ens=1, phidb=0.8, chi=15.0)


def _lamella(D0):
    x = np.arange(32) * D0 / 32
    return np.stack([2.0 * np.cos(2 * np.pi * x / D0), np.zeros(32)])


def _saddle(D):
    """Oracle-converged lamella at box D: fields W (2, M), last props r, grid g."""
    c = LAM
    g = orc.Grid(c["m"], D)
    W = _lamella(D[0])
    for _ in range(20000):
        r = g.props(W, c["N"], c["alpha"], c["phidb"], c["f"], c["ens"])
        dm = c["chi"] * (r["phi"][1] - r["phi"][0]) - 2 * W[0]
        dp = c["chi"] * (r["phi"][0] + r["phi"][1] - 1)
        if max(np.abs(dm).max(), np.abs(dp).max()) < 1e-9:
            break
        W[0] += 0.1 * dm
        W[1] += 0.05 * dp
    return W, r, g


def test_box_stress_at_saddle_matches_oracle():
    """scft_box_stress at oracle-converged fields: -d lnQ/dD and <phiA phiB> at 1e-10."""
    c = LAM
    D = (4.0, 1.0, 1.0)
    W, r, g = _saddle(D)
    _, dl = g.dlnq_dbox(W, c["N"], c["alpha"], c["phidb"], c["f"], c["ens"])
    with Context(c["m"], D, c["N"], c["f"], c["alpha"], c["ens"], c["phidb"], c["chi"], dostring=0) as ctx:
        ctx.set_fields(W[None])
        dFdD, dFdchi = ctx.box_stress()
    _say(f"saddle stress gpu {dFdD[0]} oracle {-dl}")
    assert _rel(dFdD[0], -dl) <= 1e-10
    ref = np.mean(r["phi"][0] * r["phi"][1])
    assert abs(dFdchi[0] - ref) <= 1e-10 * ref


def test_relax_box_reaches_zero_stress():
    """Stress-free relaxation of a lamella: |dF/dD| < 1e-6 with the reference's own field
    iteration. The oracle's stress at the relaxed fields and box confirms the GPU value."""
    c = LAM
    with Context(c["m"], (4.0, 1.0, 1.0), c["N"], c["f"], c["alpha"], c["ens"], c["phidb"], c["chi"],
                 dostring=0) as ctx:
        ctx.set_fields(_lamella(4.0)[None])
        D, g, nbox = ctx.relax_box(axes=(1, 0, 0), stress_tol=1e-6, err_tol=1e-9)
        W = ctx.get_fields()[0]
    _say(f"relaxed box D = {D[0]!r} after {nbox} boxes, dF/dD = {g[0]:.3e}")
    assert abs(g[0]) < 1e-6
    _, dl = orc.Grid(c["m"], D).dlnq_dbox(W.reshape(2, -1), c["N"], c["alpha"], c["phidb"], c["f"], c["ens"])
    assert abs(-dl[0] - g[0]) <= 1e-10 * max(abs(dl[0]), 1e-3)
    assert 4.3 < D[0] < 4.7


def test_autograd_backward_of_free_energy():
    """FreeEnergy.apply(ctx, D, chi) and its backward (scft_backward): forward gives FreeE's
    FEs; D.grad and chi.grad are the dL/dF-weighted sums of scft_box_stress over replicas.
    Two replicas: the saddle and its translate by a quarter period (also a saddle, same F,
    same stress). The stress itself is pinned against the oracle above."""
    torch = pytest.importorskip("torch")
    from scft_string_amd.autograd import FreeEnergy

    c = LAM
    D = (4.0, 1.0, 1.0)
    W, r, g = _saddle(D)
    W2 = np.stack([np.roll(W[0], 8), np.roll(W[1], 8)])
    _, dl = g.dlnq_dbox(W, c["N"], c["alpha"], c["phidb"], c["f"], c["ens"])
    wts = (0.3, 1.7)
    with Context(c["m"], D, c["N"], c["f"], c["alpha"], c["ens"], c["phidb"], c["chi"], nrep=2,
                 dostring=0) as ctx:
        ctx.set_fields(np.stack([W, W2]))
        dFdD, dFdchi = ctx.box_stress()
        Dt = torch.tensor(D, dtype=torch.float64, requires_grad=True)
        chit = torch.tensor(c["chi"], dtype=torch.float64, requires_grad=True)
        F = FreeEnergy.apply(ctx, Dt, chit)
        (torch.tensor(wts, dtype=torch.float64) * F).sum().backward()
        fe = ctx.free_energy()["FEs"]
        dD1, dchi1 = ctx.backward(np.array([1.0, 0.0]))
        with pytest.raises(ValueError):
            FreeEnergy.apply(ctx, Dt.detach() * 1.01, chit)
    Fh = F.detach().numpy()
    _say(f"autograd: F {Fh}, D.grad {Dt.grad.numpy()}, chi.grad {chit.grad.item()!r}")
    assert np.allclose(Fh, fe, rtol=1e-13, atol=0)
    assert abs(Fh[0] - Fh[1]) <= 1e-10 * abs(Fh[0])
    expD = wts[0] * dFdD[0] + wts[1] * dFdD[1]
    assert _rel(Dt.grad.numpy(), expD) <= 1e-12
    expchi = wts[0] * dFdchi[0] + wts[1] * dFdchi[1]
    assert abs(chit.grad.item() - expchi) <= 1e-12 * abs(expchi)
    # free_energy, like FreeE, first relaxes W+ by solve_field0 (scft.cu:817-818): at this
    # saddle that moves the fields, and the stress after it, by rounding only
    assert _rel(dD1[0], dFdD[0]) <= 1e-10 and np.all(dD1[1] == 0) and dchi1[1] == 0
    # sum_r dF_r/dD_x: both replicas carry the oracle's stress -d lnQ/dD_x
    assert _rel(Dt.grad.numpy()[0], -(wts[0] + wts[1]) * dl[0]) <= 1e-10


@pytest.mark.parametrize("D0", [5.6, 6.0, 6.4])
def test_gradient_at_64cube_N100(D0):
    """The adjoint workload's own size: 64^3, N = 100, chiN = 20, f = 0.5, a lamella-to-cylinder
    field, D swept over 5.6-6.4. FP64 d lnQ/dD against the oracle (1e-10) and against central
    differences with h = 1e-4 (1e-6: the difference quotient's own error is O(h^2));
    checkpoints every 10 contour steps against the full store (1e-12); FP32 against FP64
    (1e-4)."""
    m, N, f, alpha, ens, phidb = (64, 64, 64), 100, 0.5, 0.1, 1, 0.8
    D = (D0, D0, D0)
    M = int(np.prod(m))
    X, Y = np.meshgrid(2 * np.pi * np.arange(m[0]) / m[0], 2 * np.pi * np.arange(m[1]) / m[1], indexing="ij")
    lam = np.cos(X)
    cyl = (np.cos(X) + np.cos(X / 2 + np.sqrt(3) * Y / 2) + np.cos(X / 2 - np.sqrt(3) * Y / 2)) / 3
    W = np.zeros((1, 2, M))
    W[0, 0] = np.repeat((0.6 * lam + 0.4 * cyl)[:, :, None] * 1.5, m[2], axis=2).reshape(-1)
    res = {}
    for prec, ck in (("fp64", 0), ("fp64", 10), ("fp32", 0)):
        with Context(m, D, N, f, alpha, ens, phidb, chi=20.0, precision=prec, dostring=0, ckpt=ck) as ctx:
            lq, dq = ctx.dlnq_dbox(W)
            res[(prec, ck)] = (lq[0], dq[0])
            if prec == "fp64" and ck == 0:
                h, fd = 1e-4, np.zeros(3)
                for i in range(3):
                    lp = []
                    for s in (1, -1):
                        Dh = list(D)
                        Dh[i] += s * h
                        ctx.set_box(Dh)
                        lp.append(ctx.props_batch(W)[1][0])
                    fd[i] = (lp[0] - lp[1]) / (2 * h)
                ctx.set_box(D)
    orc.set_threads(16)
    _, dl_o = orc.Grid(m, D).dlnq_dbox(W[0], N, alpha, phidb, f, ens)
    g64, g64c, g32 = res[("fp64", 0)][1], res[("fp64", 10)][1], res[("fp32", 0)][1]
    _say(f"D={D0}: gpu {g64}, oracle {dl_o}, fd {fd}; ckpt10 rel {_rel(g64c, g64):.1e}, fp32 rel {_rel(g32, g64):.1e}")
    assert _rel(g64, dl_o) <= 1e-10
    assert _rel(g64, fd) <= 1e-6
    assert _rel(g64c, g64) <= 1e-12
    assert _rel(g32, g64) <= 1e-4


def _string_64(R, D0):
    """lamella -> cylinder string at 64^3 (the adjoint workload's fields), R replicas"""
    m = (64, 64, 64)
    M = int(np.prod(m))
    X, Y = np.meshgrid(2 * np.pi * np.arange(m[0]) / m[0], 2 * np.pi * np.arange(m[1]) / m[1], indexing="ij")
    lam = np.cos(X)
    cyl = (np.cos(X) + np.cos(X / 2 + np.sqrt(3) * Y / 2) + np.cos(X / 2 - np.sqrt(3) * Y / 2)) / 3
    W = np.zeros((R, 2, M))
    for i in range(R):
        t = i / (R - 1)
        W[i, 0] = np.repeat((((1 - t) * lam + t * cyl) * 1.5)[:, :, None], m[2], axis=2).reshape(-1)
    return m, W


def test_string_gradient_8_replicas_64cube_N100():
    """The 8-replica string of the adjoint workload (64^3, N = 100, chiN = 20, f = 0.5,
    lamella -> cylinder, D = 6.0) in one batched call: FP64 d lnQ/dD of replicas 0, 3 and 7
    against the oracle (1e-10), every replica of the batch against that replica alone
    (1e-12: batching changes nothing), and FP32 against FP64 (1e-4)."""
    R, D0, N, f, alpha, ens, phidb = 8, 6.0, 100, 0.5, 0.1, 1, 0.8
    m, W = _string_64(R, D0)
    D = (D0, D0, D0)
    res = {}
    for prec in ("fp64", "fp32"):
        with Context(m, D, N, f, alpha, ens, phidb, chi=20.0, nrep=R, precision=prec, fix=(1, 1)) as ctx:
            res[prec] = ctx.dlnq_dbox(W)[1]
            if prec == "fp64":
                single = ctx.dlnq_dbox(W[3:4])[1][0]
    orc.set_threads(16)
    g = orc.Grid(m, D)
    for i in (0, 3, 7):
        _, dl = g.dlnq_dbox(W[i], N, alpha, phidb, f, ens)
        _say(f"string replica {i}: gpu {res['fp64'][i]}, oracle {dl}")
        assert _rel(res["fp64"][i], dl) <= 1e-10
    assert _rel(single, res["fp64"][3]) <= 1e-12
    worst = max(_rel(res["fp32"][i], res["fp64"][i]) for i in range(R))
    _say(f"8-replica string: FP32 against FP64 worst rel {worst:.1e}")
    assert worst <= 1e-4


def test_relax_box_64cube_N100_to_zero_stress():
    """A stress-free relaxation at the adjoint workload's size: a lamella along x in a 64^3
    box (chiN = 20, f = 0.5, phidb = 0.8, N = 100, canonical, FP64), box length x relaxed
    to |dF/dD_x| < 1e-6, the fields converged at every box by Anderson mixing. The oracle's
    stress at the final fields and box confirms the value."""
    m, N, f, alpha, ens, phidb, chi = (64, 64, 64), 100, 0.5, 0.1, 1, 0.8, 20.0
    D0 = (4.2, 4.0, 4.0)
    M = int(np.prod(m))
    x = (np.arange(m[0]) * D0[0] / m[0])[:, None, None] * np.ones(m)
    W0 = np.zeros((1, 2, M))
    W0[0, 0] = (2.0 * np.cos(2 * np.pi * x / D0[0])).reshape(-1)
    with Context(m, D0, N, f, alpha, ens, phidb, chi, dostring=0, anderson=20) as ctx:
        ctx.set_fields(W0)
        D, gst, nbox = ctx.relax_box(axes=(1, 0, 0), stress_tol=1e-6, err_tol=1e-9, max_field_iter=4000,
                                     max_boxes=30)
        W = ctx.get_fields()[0]
    _say(f"64^3 lamella relaxed: D = {D[0]!r} after {nbox} boxes, dF/dD = {gst}")
    assert abs(gst[0]) < 1e-6
    orc.set_threads(16)
    _, dl = orc.Grid(m, D).dlnq_dbox(W.reshape(2, -1), N, alpha, phidb, f, ens)
    assert abs(-dl[0] - gst[0]) <= 1e-9
    assert abs(dl[1]) < 1e-9 and abs(dl[2]) < 1e-9   # uniform along y and z: no stress there
