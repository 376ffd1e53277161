"""Pins the FP64 CPU oracle (oracle/scft_oracle.c) before it is trusted as the checker.

Every oracle function is checked against something it does not share code with:
numpy/scipy golden fixtures, closed forms, or invariants that a dropped term, a wrong
sign or a swapped species would break.
"""
import math
import os

import numpy as np
import pytest

from oracle import orc

GOLD = np.load(os.path.join(os.path.dirname(__file__), "golden", "fixtures.npz"))


# ---------------------------------------------------------------- FFT (cuFFT D2Z / Z2D)
@pytest.mark.parametrize("i", range(int(GOLD["nfft"])))
def test_rfft3_matches_numpy(i):
    shp = tuple(int(v) for v in GOLD[f"fft{i}_shape"])
    g = orc.Grid(shp, (1.0, 1.0, 1.0))
    got = g.rfft3(GOLD[f"fft{i}_in"])
    ref = GOLD[f"fft{i}_rfft"]
    assert np.max(np.abs(got - ref)) <= 1e-12 * np.max(np.abs(ref))


@pytest.mark.parametrize("i", range(int(GOLD["nfft"])))
def test_irfft3_matches_numpy(i):
    shp = tuple(int(v) for v in GOLD[f"fft{i}_shape"])
    g = orc.Grid(shp, (1.0, 1.0, 1.0))
    got = g.irfft3(GOLD[f"fft{i}_hin"])
    ref = GOLD[f"fft{i}_irfft"]
    assert np.max(np.abs(got - ref)) <= 1e-12 * np.max(np.abs(ref))


# ---------------------------------------------------------------- A1 k^2 table
@pytest.mark.parametrize("m,D", [((8, 8, 8), (4.0, 4.0, 4.0)), ((6, 10, 12), (3.0, 5.0, 2.5)),
                                 ((5, 7, 9), (1.0, 2.0, 3.0))])
def test_ksq_closed_form(m, D):
    g = orc.Grid(m, D)
    ks = g.ksq()
    f0 = np.fft.fftfreq(m[0]) * m[0]
    f1 = np.fft.fftfreq(m[1]) * m[1]
    f2 = np.arange(m[2] // 2 + 1)
    ref = 4 * math.pi ** 2 * (f0[:, None, None] ** 2 / D[0] ** 2 + f1[None, :, None] ** 2 / D[1] ** 2
                              + f2[None, None, :] ** 2 / D[2] ** 2)
    np.testing.assert_allclose(ks, ref, rtol=1e-14, atol=1e-12)


# ---------------------------------------------------------------- A4 step closed forms
def _factors(g, w, ds):
    M = g.M
    k12 = np.exp(-g.ksq().reshape(-1) * ds / 12) / M
    k6 = k12 * k12 * M
    e4 = np.exp(-np.asarray(w).reshape(-1) * ds / 4)
    return e4, e4 * e4, k12, k6


def test_step_uniform_field_is_exact_exponential():
    g = orc.Grid((8, 6, 10), (3.0, 4.0, 5.0))
    ds = 1 / 50
    w = np.full(g.M, 1.7)
    q = np.full(g.M, 0.3)
    out = g.step(*_factors(g, w, ds), q)
    np.testing.assert_allclose(out.reshape(-1), 0.3 * np.exp(-1.7 * ds), rtol=1e-13)


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_step_single_mode_decays_as_kinetic_exponential(axis):
    m, D = (12, 8, 10), (3.0, 2.0, 2.5)
    g = orc.Grid(m, D)
    ds = 1 / 40
    n = 2
    coords = np.meshgrid(*[np.arange(mi) * D[i] / mi for i, mi in enumerate(m)], indexing="ij")
    q = np.cos(2 * np.pi * n * coords[axis] / D[axis])
    out = g.step(*_factors(g, np.zeros(g.M), ds), q)
    k2 = (2 * np.pi * n / D[axis]) ** 2
    np.testing.assert_allclose(out, q * np.exp(-k2 * ds / 6), atol=1e-13)


def test_step_is_not_plain_split_step():
    """Richardson weights: output differs from either half- or full-step alone."""
    g = orc.Grid((8, 8, 8), (2.0, 2.0, 2.0))
    rng = np.random.default_rng(1)
    w = rng.standard_normal(g.M) * 3
    ds = 1 / 20
    e4, e2, k12, k6 = _factors(g, w, ds)
    q = np.ones(g.M)
    out = g.step(e4, e2, k12, k6, q).reshape(-1)
    # brute-force numpy restatement of one full split step, for contrast
    def K(a, kf):
        return np.fft.irfftn(np.fft.rfftn(a.reshape(g.m)) * kf.reshape(8, 8, 5), s=g.m,
                             axes=(0, 1, 2)).reshape(-1) * g.M
    full = e2 * K(e2 * q, k6)
    half = e4 * K(e2 * K(e4 * q, k12), k12)
    np.testing.assert_allclose(out, (4 * half - full) / 3, rtol=1e-12)
    assert np.max(np.abs(out - full)) > 1e-6


# ---------------------------------------------------------------- A5-A9 props
def _random_W(M, seed, amp=2.0):
    rng = np.random.default_rng(seed)
    return np.stack([amp * rng.standard_normal(M), 0.5 * rng.standard_normal(M)])


def test_props_symmetric_operator_invariant():
    """<q1[s] q2[N-s]> = QC for every s; <q2[s] q2[Nh-s]> = Q2 (A4 symmetry)."""
    g = orc.Grid((8, 6, 10), (3.0, 3.0, 3.0))
    N = 24
    W = _random_W(g.M, 3)
    out = g.props(W, N=N, alpha=0.25, phic=0.7, f=0.4, ens=1, want_q=True)
    q = out["q"]
    QC = out["Q"][2]
    Q2 = out["Q"][1]
    Nh = int(0.25 * N)
    for s in range(N + 1):
        assert abs(np.mean(q[s, 0] * q[N - s, 1]) - QC) <= 1e-12 * QC
    for s in range(Nh + 1):
        assert abs(np.mean(q[s, 1] * q[Nh - s, 1]) - Q2) <= 1e-12 * Q2
    # the species switch is real: q1 differs from q2 at s = N
    assert np.max(np.abs(q[N, 0] - q[N, 1])) > 1e-3


@pytest.mark.parametrize("N,f,alpha", [(100, 0.5, 0.1), (128, 0.453125, 0.1), (128, 0.3984375, 0.1),
                                       (64, 0.40625, 0.1)])
def test_props_W0_simpson_closed_forms_canonical(N, f, alpha):
    g = orc.Grid((4, 4, 4), (2.0, 2.0, 2.0))
    phic = 0.8
    out = g.props(np.zeros((2, g.M)), N=N, alpha=alpha, phic=phic, f=f, ens=1)
    n1 = int(round(N * f))
    n2 = N - n1
    Nh = int(alpha * N)

    def wsum(n):  # 2 end terms + interior weights 4,2,4,... over n-1 points (density.h:405-422)
        return 2 + sum(4 if i % 2 == 0 else 2 for i in range(n - 1))

    ds = 1.0 / N
    phiA, phiB, phih, phicB = out["phi"]
    np.testing.assert_allclose(phiA, phic * ds / 3 * wsum(n1), rtol=1e-13)
    np.testing.assert_allclose(phicB, phic * ds / 3 * wsum(n2), rtol=1e-13)
    np.testing.assert_allclose(phih, (1 - phic) / alpha * ds / 3 * wsum(Nh), rtol=1e-13)
    np.testing.assert_allclose(phiB, phih + phicB, rtol=1e-14)
    assert out["lnQ"] == 0.0
    # odd-count quirk: sum of weights is 3n-1 for odd n, 3n for even n (A.8)
    for n in (n1, n2, Nh):
        assert wsum(n) == (3 * n if n % 2 == 0 else 3 * n - 1)
    np.testing.assert_allclose(out["phis"], [np.mean(phiA + phicB), np.mean(phiA), np.mean(phiB), 0.0,
                                             np.mean(phih)], rtol=1e-13)


def test_props_W0_grand_canonical():
    g = orc.Grid((4, 4, 4), (2.0, 2.0, 2.0))
    out = g.props(np.zeros((2, g.M)), N=60, alpha=0.1, phic=0.3, f=0.8, ens=2)
    assert abs(out["lnQ"] - (1.0 + 0.3 * 1.0)) < 1e-14
    np.testing.assert_allclose(out["phi"][0], 0.3 * (1 / 60) / 3 * 3 * 48, rtol=1e-13)


def test_props_uniform_field_Q_closed_form():
    g = orc.Grid((4, 6, 4), (2.0, 2.0, 2.0))
    a, b = 0.7, -1.3  # W-, W+
    W = np.stack([np.full(g.M, a), np.full(g.M, b)])
    N, f, alpha = 50, 0.42, 0.1
    out = g.props(W, N=N, alpha=alpha, phic=0.6, f=f, ens=1)
    n1 = int(round(N * f))
    n2 = N - n1
    Nh = int(alpha * N)
    wA, wB = a + b, b - a
    np.testing.assert_allclose(out["Q"][2], np.exp(-(n1 * wA + n2 * wB) / N), rtol=1e-12)
    np.testing.assert_allclose(out["Q"][1], np.exp(-Nh * wB / N), rtol=1e-12)


def test_props_canonical_normalisation_random_field():
    """For even n1, n2 the copolymer density integrates to phic (uses A4 symmetry)."""
    g = orc.Grid((8, 8, 6), (3.0, 3.0, 2.0))
    W = _random_W(g.M, 7)
    out = g.props(W, N=40, alpha=0.1, phic=0.65, f=0.5, ens=1)
    phiA, phiB, phih, phicB = out["phi"]
    assert abs(np.mean(phiA + phicB) - 0.65) < 1e-12
    assert abs(np.mean(phih) - 0.35) < 1e-12  # Nh = 4 even and Nh = alpha N


# ---------------------------------------------------------------- A13 spline
@pytest.mark.parametrize("i", range(int(GOLD["nspl"])))
def test_spline_matches_scipy_natural(i):
    x, y, xr = GOLD[f"spl{i}_x"], GOLD[f"spl{i}_y"], GOLD[f"spl{i}_xref"]
    d = orc.spline_deriv_nodes(x, y)
    v = orc.spline_value_at(x, y, xr)
    np.testing.assert_allclose(d, GOLD[f"spl{i}_deriv"], rtol=1e-10, atol=1e-10)
    np.testing.assert_allclose(v, GOLD[f"spl{i}_value"], rtol=1e-10, atol=1e-11)


def test_spline_linear_data_exact():
    x = np.array([0.0, 0.1, 0.35, 0.6, 0.61, 1.0])
    y = 2.5 * x - 1.0
    np.testing.assert_allclose(orc.spline_deriv_nodes(x, y), 2.5, rtol=1e-12)
    xr = np.arange(6) / 5.0
    np.testing.assert_allclose(orc.spline_value_at(x, y, xr), 2.5 * xr - 1.0, atol=1e-13)


# ---------------------------------------------------------------- A10 solve_field0
def test_solve_field0_uniform_drift_never_converges():
    """Canonical, W = 0, Nh != alpha*N: DEV+ has a mean W+ cannot remove (A.3).
    W+ stays uniform, W- is untouched and all 19 sub-iterations run."""
    g = orc.Grid((4, 4, 4), (2.0, 2.0, 2.0))
    M = g.M
    N, alpha, phic, chi, f = 128, 0.1, 0.9, 18.0, 0.453125
    row = np.zeros(7 * M)
    dev = np.zeros(2 * M)
    k = g.solve_field0(row, dev, chi, f, N, alpha, phic, 1, 20, 1e-3)
    assert k == 20
    n1 = int(round(N * f))
    Nh = int(alpha * N)

    def wsum(n):
        return 2 + sum(4 if i % 2 == 0 else 2 for i in range(n - 1))

    tot = phic / (3 * N) * (wsum(n1) + wsum(N - n1)) + (1 - phic) / alpha / (3 * N) * wsum(Nh)
    d = chi * (tot - 1.0)
    np.testing.assert_allclose(dev[M:], d, rtol=1e-12)  # last DEV+ is the same constant
    wp = row[M:2 * M]
    assert np.ptp(wp) <= 1e-14 * abs(wp[0])
    assert 19 * 0.005 * abs(d) <= abs(wp[0]) <= 19 * 0.1 * abs(d) and wp[0] * d > 0
    np.testing.assert_array_equal(row[:M], 0.0)


def test_solve_field0_grand_canonical_closed_form():
    """GC, W- = 0, uniform W+ = c: phiA+phiB = z_C e^{-c} + alpha e^{-c alpha} exactly
    (even n1, n2, Nh = alpha N), so the whole 19-step mixing map has a closed form."""
    g = orc.Grid((4, 4, 4), (2.0, 2.0, 2.0))
    M = g.M
    N, alpha, zc, chi, f = 100, 0.1, 0.5, 10.0, 0.5
    row = np.zeros(7 * M)
    dev = np.zeros(2 * M)
    k = g.solve_field0(row, dev, chi, f, N, alpha, zc, 2, 20, 1e-9)
    lam, errt, c = 0.05, 1.0, 0.0
    kk = 1
    err = 1.0
    while kk < 20 and err > 1e-9:
        d = chi * (zc * math.exp(-c) + alpha * math.exp(-c * alpha) - 1.0)
        err = abs(d)
        c += lam * d
        lam = min(0.1, lam * 1.09) if err < errt else max(0.005, lam / 1.10)
        errt = err
        kk += 1
    assert k == kk
    np.testing.assert_allclose(row[M:2 * M], c, rtol=1e-12)


def test_solve_field0_converged_exits_at_k2():
    g = orc.Grid((4, 4, 4), (2.0, 2.0, 2.0))
    M = g.M
    row = np.zeros(7 * M)
    dev = np.zeros(2 * M)
    k = g.solve_field0(row, dev, 20.0, 0.5, 100, 0.1, 0.8, 1, 5000, 1e-4)
    assert k == 2


# ---------------------------------------------------------------- A16 FreeE closed form
def test_free_energy_disordered_closed_form():
    g = orc.Grid((4, 4, 4), (2.0, 2.0, 2.0))
    M = g.M
    R = 3
    N, f, phic, chi = 100, 0.5, 0.8, 20.0
    st = orc.StringState(g, dict(chi=chi, f=f, alpha=0.1, phidb=phic, N=N, ens=1, dostring=1, fix=(1, 1)),
                         np.zeros((R, 2, M)))
    fe = st.free_energy()
    phiA = phic * f
    phiB = phic * (1 - f) + (1 - phic)
    np.testing.assert_allclose(fe["FEs"], chi * phiA * phiB, rtol=1e-12)
    np.testing.assert_allclose(fe["phctot"], phic, rtol=1e-12)
    assert abs(fe["FE"] - R * chi * phiA * phiB) < 1e-10


# ---------------------------------------------------------------- A11-A15 string iteration
def test_string_iteration_projection_and_arclength():
    g = orc.Grid((6, 4, 4), (3.0, 2.0, 2.0))
    M = g.M
    R = 5
    x = np.arange(6)[:, None, None] * 3.0 / 6
    base = np.cos(2 * np.pi * x / 3.0) * np.ones((6, 4, 4))
    rng = np.random.default_rng(5)
    W0 = np.zeros((R, 2, M))
    for i in range(R):
        W0[i, 0] = (0.5 + i * 0.4) * base.reshape(-1) + 0.05 * rng.standard_normal(M)
    st = orc.StringState(g, dict(chi=14.0, f=0.5, alpha=0.1, phidb=0.8, N=20, ens=1, dostring=1,
                                 fix=(1, 1)), W0)
    Wm_before = st.W[:, :M].copy()
    rep = st.iterate(0)
    # arc length from W- only, normalised, ends pinned
    assert rep["x"][0] == 0 and rep["x"][-1] == 1
    assert np.all(np.diff(rep["x"]) > 0)
    # projection removes the tangential part for interior replicas
    assert np.all(np.abs(rep["dotWs2"][1:-1]) < 1e-12)
    assert np.all(np.abs(rep["dotWs"][1:-1]) > 1e-8)
    # fixed ends are never propped or moved
    assert rep["steps"][0] == 0 and rep["steps"][-1] == 0
    np.testing.assert_array_equal(st.W[0, :M], Wm_before[0])
    np.testing.assert_array_equal(st.W[-1, :M], Wm_before[-1])
    assert all(1 <= s <= 20 for s in rep["steps"][1:-1])
