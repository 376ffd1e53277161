"""Pins of the oracle's Anderson mixing (orc_anderson_iter), the opt-in mode for dostring = 0
that the reference declares arrays for but never runs (scft.cu:30-31, 1029-1031).

- H = 0 with the Anderson phase from the start is simple mixing W += lambda DEV.
- The converged fields are the saddle point that simple mixing reaches.
- It converges in far fewer iterations than simple mixing.
- The small elimination solve equals numpy.linalg.solve.
"""
import ctypes as C

import numpy as np
import pytest

from oracle import orc

M3, D3, N, CHI, F, ALPHA, PHIC, ENS = (32, 1, 1), (4.0, 1.0, 1.0), 40, 15.0, 0.5, 0.1, 0.8, 1


def _lamella():
    x = np.arange(M3[0]) * D3[0] / M3[0]
    return np.stack([2.0 * np.cos(2 * np.pi * x / D3[0]), np.zeros(M3[0])])


def _simple_step(g, W, lam):
    r = g.props(W, N, ALPHA, PHIC, F, ENS)
    pA, pB = r["phi"][0], r["phi"][1]
    d = np.concatenate([CHI * (pB - pA) - 2 * W[0], CHI * (pA + pB - 1)])
    W += lam * d.reshape(2, -1)
    return np.sqrt(np.mean(d * d))


@pytest.fixture(autouse=True)
def _one_thread():
    orc.set_threads(1)
    yield
    import os
    orc.set_threads(os.cpu_count() or 1)


def test_no_history_is_simple_mixing():
    g = orc.Grid(M3, D3)
    Wa, Ws = _lamella(), _lamella()
    a = orc.Anderson(g, 0, CHI, F, N, ALPHA, PHIC, ENS, 0.05, 1e30)
    for _ in range(20):
        ea = a.iterate(Wa)
        es = _simple_step(g, Ws, 0.05)
        assert abs(ea - es) <= 1e-14 * es
    assert np.max(np.abs(Wa - Ws)) <= 1e-13


def test_same_saddle_point_and_faster_than_simple_mixing():
    """Started at errs < 1e-2 (the default), Anderson reaches simple mixing's lamella in ~4.6x
    fewer iterations. Started at errs < 0.1 it lands on the disordered saddle point
    (W- = chi (phiB - phiA)/2 = 1.5 everywhere), which is why the default is 1e-2."""
    g = orc.Grid(M3, D3)
    Ws = _lamella()
    ks = next(k for k in range(20000) if _simple_step(g, Ws, 0.05) < 1e-10)
    Wa = _lamella()
    a = orc.Anderson(g, 20, CHI, F, N, ALPHA, PHIC, ENS, 0.2, 1e-2)
    ka = next(k for k in range(5000) if a.iterate(Wa) < 1e-10)
    assert ka * 4 < ks, (ka, ks)
    assert np.max(np.abs(Wa - Ws)) <= 1e-8
    Wd = _lamella()
    d = orc.Anderson(g, 20, CHI, F, N, ALPHA, PHIC, ENS, 0.2, 0.1)
    next(k for k in range(5000) if d.iterate(Wd) < 1e-10)
    assert np.max(np.abs(Wd[0] - 1.5)) <= 1e-8


def test_small_solve_matches_numpy():
    rng = np.random.default_rng(1)
    for n in (1, 3, 7, 15):
        U = rng.standard_normal((n, n)) + n * np.eye(n)
        V = rng.standard_normal(n)
        Uc, Vc, c = U.copy().reshape(-1), V.copy(), np.zeros(n)
        dp = C.POINTER(C.c_double)
        ok = orc.lib().orc_small_solve(n, Uc.ctypes.data_as(dp), Vc.ctypes.data_as(dp), c.ctypes.data_as(dp))
        assert ok == 1
        assert np.max(np.abs(c - np.linalg.solve(U, V))) <= 1e-12 * np.max(np.abs(c))
    Us = np.ones(4)
    Vs = np.ones(2)
    c = np.zeros(2)
    dp = C.POINTER(C.c_double)
    assert orc.lib().orc_small_solve(2, Us.ctypes.data_as(dp), Vs.ctypes.data_as(dp), c.ctypes.data_as(dp)) == 0
