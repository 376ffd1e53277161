"""Pins of the oracle's A homopolymer (orc_grid_set_phiha), which the reference declares but
holds at zA = 0 (density.h:352, 390-397; README.md:3).

The A <-> B mirror: negate W- (wA <-> wB) and swap f <-> 1 - f. Read from the other end,
the copolymer is then the same chain, and an A homopolymer becomes a B homopolymer of the
same length. So a blend with an A homopolymer must reproduce, with phiA and phiB swapped,
the reference's own B-homopolymer blend. The Simpson weights restart at each block's
first interior point, so the mirror is exact only for even n1, n2 and Nh.
"""
import numpy as np
import pytest

from oracle import orc


def _fields(m, D, seed):
    rng = np.random.default_rng(seed)
    M = int(np.prod(m))
    x = (np.arange(m[0]) * D[0] / m[0])[:, None, None] * np.ones(m)
    Wm = 1.5 * np.cos(2 * np.pi * x / D[0]).reshape(-1) + 0.3 * rng.standard_normal(M)
    Wp = 0.2 * rng.standard_normal(M)
    return np.stack([Wm, Wp])


M3, D3, N, ALPHA = (8, 6, 10), (2.0, 1.5, 2.5), 20, 0.1   # n1 = 8, n2 = 12, Nh = 2


def _props(W, f, ens, phic, phihA):
    g = orc.Grid(M3, D3)
    g.set_phihA(phihA)
    return g.props(W, N, ALPHA, phic, f, ens)


@pytest.mark.parametrize("ens,phic,phihA_a,phihA_b", [(1, 0.7, 0.3, 0.0), (2, 0.5, 1.0, 1.0)])
def test_mirror_maps_a_homopolymer_onto_b_homopolymer(ens, phic, phihA_a, phihA_b):
    """ens 1: A-homopolymer blend (phihB = 0) vs the reference's B blend (phihA = 0).
    ens 2: both homopolymers at activity 1 on both sides (the reference fixes zB = 1)."""
    W = _fields(M3, D3, 4)
    Wm = np.stack([-W[0], W[1]])
    a = _props(W, 0.4, ens, phic, phihA_a)
    b = _props(Wm, 0.6, ens, phic, phihA_b)
    assert abs(a["lnQ"] - b["lnQ"]) <= 1e-13 * abs(b["lnQ"])
    for ia, ib in [(0, 1), (1, 0), (2, 2)]:   # phiA <-> phiB, phih total
        assert np.max(np.abs(a["phi"][ia] - b["phi"][ib])) <= 1e-13, (ia, ib)
    # the copolymer B-block density phicB maps onto the mirror's copolymer A block, phiA - phihA
    hA_b = b["phis"][3]  # the mirror has only a B homopolymer in ens 1: phihA = 0 there
    assert np.max(np.abs(a["phi"][3] - (b["phi"][0] - (b["phi"][2] - b["phi"][1] + b["phi"][3])))) <= 1e-13
    assert hA_b == 0.0 or ens == 2


def test_uniform_field_normalisation():
    """W = 0: canonical densities are uniform and each species sums to its fraction."""
    M = int(np.prod(M3))
    W = np.zeros((2, M))
    r = _props(W, 0.4, 1, 0.6, 0.25)
    phiA, phiB, phih = r["phi"][0], r["phi"][1], r["phi"][2]
    assert np.allclose(phiA + phiB, 1.0, atol=1e-13)
    assert abs(np.mean(phih) - (1.0 - 0.6)) <= 1e-13          # both homopolymers
    assert abs(np.mean(phiA) - (0.6 * 0.4 + 0.25)) <= 1e-13   # copolymer A block + A homopolymer
    assert abs(r["phis"][3] - 0.25) <= 1e-13                   # <phi_hA>
    assert abs(r["phis"][4] - 0.15) <= 1e-13                   # <phi_hB>


@pytest.mark.parametrize("ens,phic,phihA", [(1, 0.6, 0.25), (2, 0.4, 0.7)])
def test_dlnq_dbox_with_a_homopolymer_matches_central_differences(ens, phic, phihA):
    W = _fields(M3, D3, 6)
    g = orc.Grid(M3, D3)
    g.set_phihA(phihA)
    _, d = g.dlnq_dbox(W, N, ALPHA, phic, 0.4, ens)
    fd = np.zeros(3)
    for i in range(3):
        h = 1e-3 * D3[i]
        v = []
        for k in (2, 1, -1, -2):
            Dk = list(D3)
            Dk[i] += k * h
            gk = orc.Grid(M3, Dk)
            gk.set_phihA(phihA)
            v.append(gk.props(W, N, ALPHA, phic, 0.4, ens)["lnQ"])
        fd[i] = (-v[0] + 8 * v[1] - 8 * v[2] + v[3]) / (12 * h)
    assert np.max(np.abs(d - fd)) <= 1e-9 * np.max(np.abs(fd)), (d, fd)
