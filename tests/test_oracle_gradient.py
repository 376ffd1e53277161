"""Pins of the oracle's box derivative orc_dlnq_dbox (the adjoint of density.h:344-399).

The oracle evaluates d lnQ / dD_i from the stored q1, q2 as adjoints, with every dK/dD
applied in real space. It is checked against 4-point central differences of orc_props
itself (error O(h^4), h = 1e-3 D). A uniform field gives exactly zero, because only the
k = 0 mode is populated and its kinetic factor does not depend on D.
"""
import os

import numpy as np
import pytest

from oracle import orc


def _fields(m, D, seed):
    rng = np.random.default_rng(seed)
    M = int(np.prod(m))
    x = (np.arange(m[0]) * D[0] / m[0])[:, None, None] * np.ones(m)
    y = (np.arange(m[1]) * D[1] / m[1])[None, :, None] * np.ones(m)
    Wm = (2.0 * np.cos(2 * np.pi * x / D[0]) + np.cos(2 * np.pi * y / D[1])).reshape(-1)
    Wm += 0.3 * rng.standard_normal(M)
    Wp = 0.2 * rng.standard_normal(M)
    return np.stack([Wm, Wp])


def _fd(m, D, W, N, alpha, phic, f, ens):
    out = np.zeros(3)
    for i in range(3):
        h = 1e-3 * D[i]
        v = []
        for k in (2, 1, -1, -2):
            Dk = list(D)
            Dk[i] += k * h
            v.append(orc.Grid(m, Dk).props(W, N, alpha, phic, f, ens)["lnQ"])
        out[i] = (-v[0] + 8 * v[1] - 8 * v[2] + v[3]) / (12 * h)
    return out


CASES = [
    # m, D, N, f, alpha, ens, phic
    ((8, 12, 10), (2.0, 3.0, 2.5), 20, 0.5, 0.1, 1, 0.8),
    ((12, 8, 16), (3.1, 2.2, 3.7), 21, 0.4, 0.1, 2, 0.3),
    ((10, 9, 6), (2.5, 2.0, 1.5), 17, 0.55, 0.2, 1, 0.6),   # odd n1, Nh = 3 != alpha N
    ((16, 1, 1), (3.0, 1.0, 1.0), 30, 0.45, 0.1, 1, 0.9),   # 1-D box
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: "x".join(map(str, c[0])) + f"_ens{c[5]}")
def test_dlnq_dbox_matches_central_differences(case):
    m, D, N, f, alpha, ens, phic = case
    W = _fields(m, D, 3)
    lnQ, d = orc.Grid(m, D).dlnq_dbox(W, N, alpha, phic, f, ens)
    fd = _fd(m, D, W, N, alpha, phic, f, ens)
    scale = np.max(np.abs(fd))
    assert np.max(np.abs(d - fd)) <= 1e-9 * scale, (d, fd)
    ref = orc.Grid(m, D).props(W, N, alpha, phic, f, ens)["lnQ"]
    assert lnQ == ref


def test_dlnq_dbox_uniform_field_is_zero():
    m, D = (8, 6, 10), (2.0, 1.5, 2.5)
    M = int(np.prod(m))
    W = np.stack([np.full(M, 0.7), np.full(M, -0.2)])
    _, d = orc.Grid(m, D).dlnq_dbox(W, 20, 0.1, 0.8, 0.5, 1)
    assert np.max(np.abs(d)) <= 1e-14


# ---------------------------------------------------------------- at a saddle point
def _converge(g, W, chi, N, f, alpha, phic, ens, tol=1e-10):
    """Plain mixing to DEV = 0 (any solver reaches the same saddle point)."""
    for _ in range(100000):
        r = g.props(W, N, alpha, phic, f, ens)
        pA, pB = r["phi"][0], r["phi"][1]
        dm = chi * (pB - pA) - 2 * W[0]
        dp = chi * (pA + pB - 1)
        if max(np.abs(dm).max(), np.abs(dp).max()) < tol:
            return W, r
        W[0] += 0.1 * dm
        W[1] += 0.05 * dp
    raise RuntimeError("no convergence")


def _free_energy(W, r, chi):
    """FreeE's FEs = -lnQ + mean(chi phiA phiB - wA phiA - wB phiB) (scft.cu:823-830)."""
    pA, pB = r["phi"][0], r["phi"][1]
    wA, wB = W[1] + W[0], W[1] - W[0]
    return -r["lnQ"] + np.mean(chi * pA * pB - wA * pA - wB * pB)


def _envelope(N, with_chi):
    m, D0, f, alpha, phic, ens, chi = (16, 1, 1), 4.0, 0.5, 0.1, 0.8, 1, 15.0
    x = np.arange(m[0]) * D0 / m[0]
    W = np.stack([2.0 * np.cos(2 * np.pi * x / D0), np.zeros(m[0])])
    W0, r0 = _converge(orc.Grid(m, (D0, 1.0, 1.0)), W, chi, N, f, alpha, phic, ens)
    _, dl = orc.Grid(m, (D0, 1.0, 1.0)).dlnq_dbox(W0, N, alpha, phic, f, ens)

    def fd(fun, h):
        v = [fun(k * h) for k in (2, 1, -1, -2)]
        return (-v[0] + 8 * v[1] - 8 * v[2] + v[3]) / (12 * h)

    def F_at_D(dD):
        g = orc.Grid(m, (D0 + dD, 1.0, 1.0))
        Wk, rk = _converge(g, W0.copy(), chi, N, f, alpha, phic, ens)
        return _free_energy(Wk, rk, chi)

    def F_at_chi(dc):
        Wk, rk = _converge(orc.Grid(m, (D0, 1.0, 1.0)), W0.copy(), chi + dc, N, f, alpha, phic, ens)
        return _free_energy(Wk, rk, chi + dc)

    out = {"stress": -dl[0], "stress_fd": fd(F_at_D, 0.02)}
    if with_chi:
        out["dchi"] = float(np.mean(r0["phi"][0] * r0["phi"][1]))
        out["dchi_fd"] = fd(F_at_chi, 0.05)
    return out


def test_saddle_point_derivatives_match_free_energy_differences():
    """dF/dD = -d lnQ/dD and dF/dchiN = <phiA phiB> at a saddle point (the envelope theorem),
    against 4-point differences of FreeE's F with the fields re-converged at every point.
    The discrete F is stationary in W only up to the contour discretisation, so the match
    is O(ds^3): 7.5e-5 at N = 40 for 32 points, and it improves ~8x per halving of ds."""
    orc.set_threads(1)  # 16 grid points: threads only add synchronisation
    try:
        r40 = _envelope(40, True)
        r80 = _envelope(80, False)
    finally:
        orc.set_threads(os.cpu_count() or 1)
    e40 = abs(r40["stress"] - r40["stress_fd"]) / abs(r40["stress_fd"])
    c40 = abs(r40["dchi"] - r40["dchi_fd"]) / abs(r40["dchi_fd"])
    assert e40 <= 3e-4 and c40 <= 3e-4, (r40, e40, c40)
    e80 = abs(r80["stress"] - r80["stress_fd"]) / abs(r80["stress_fd"])
    assert e80 <= e40 / 4, (e40, e80)
