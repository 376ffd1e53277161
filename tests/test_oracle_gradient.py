"""Pins of the oracle's box derivative orc_dlnq_dbox (the adjoint of density.h:344-399).

The oracle evaluates d lnQ / dD_i from the stored q1, q2 as adjoints, with every dK/dD
applied in real space. It is checked against 4-point central differences of orc_props
itself (error O(h^4), h = 1e-3 D). A uniform field gives exactly zero, because only the
k = 0 mode is populated and its kinetic factor does not depend on D.
"""
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
