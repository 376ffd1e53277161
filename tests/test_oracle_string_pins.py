"""Pins of the oracle's string iteration (orc_string_iter) and FreeE by closed forms.

For spatially uniform replicas every quantity of one solve_field iteration
(scft.cu:592-757) and of FreeE (scft.cu:812-853) has a closed form
(tests/uniform_string.py). These tests hold the oracle to those forms:

- the fixed mixing parameter 0.05 (scft.cu:562, 690-693);
- errs and errs3 as RMS over 2M points (scft.cu:624-630, 673-674);
- the spring term 0.5 x the second difference in the spring phase (scft.cu:606);
- the arc length from W- only (scft.cu:351-355, 612-622): W+ differs per replica here;
- the perpendicular projection, which removes all of DEV- for the interior replicas of a
  collinear string and none for the ends (scft.cu:641-684);
- the interior-only redistribution 0.05 of the way to the spline at x = i/(R-1), from x
  recomputed after mixing (scft.cu:725-757, density.h:511-528);
- FreeE's -lnQ + <chi phiA phiB - wA phiA - wB phiB>, phctot and the cumulative W-
  distance (scft.cu:812-853).

tests/test_oracle_mutations.py rebuilds the oracle with each of these planted wrong and
checks that this file fails.
"""
import numpy as np
import pytest

from oracle import orc
from tests.uniform_string import Blend, free_energy, replay_main, string_iteration

GRID = ((4, 4, 4), (2.0, 3.0, 2.5))
R = 5
A0 = np.array([-1.2, -0.7, -0.1, 0.3, 1.4])      # W- per replica, unequally spaced
B0 = np.array([0.3, -0.2, 0.45, 0.1, -0.35])     # W+ per replica, all different
CANON = dict(N=20, f=0.5, alpha=0.1, phidb=0.8, chi=15.0, ens=1)   # n1 = n2 = 10, Nh = 2
GRAND = dict(N=20, f=0.5, alpha=0.1, phidb=0.5, chi=12.0, ens=2)


def _state(run, a, b, fix, springsteps=0):
    g = orc.Grid(*GRID)
    M = g.M
    W0 = np.zeros((R, 2, M))
    W0[:, 0] = np.asarray(a)[:, None]
    W0[:, 1] = np.asarray(b)[:, None]
    kw = dict(run, dostring=1, fix=fix, springsteps=springsteps)
    return g, orc.StringState(g, kw, W0)


def _rows(st, M):
    Wm = st.W[:, :M]
    Wp = st.W[:, M:2 * M]
    # the fields stay uniform: every point of a row equals its first point
    assert np.ptp(Wm, axis=1).max() < 1e-13 and np.ptp(Wp, axis=1).max() < 1e-13
    return Wm[:, 0], Wp[:, 0]


def _check_report(got, want, keys):
    for key in keys:
        np.testing.assert_allclose(got[key], want[key], rtol=1e-11, atol=1e-12, err_msg=key)


@pytest.mark.parametrize("fix", [(1, 1), (0, 0)])
def test_spring_phase_mixing_errs_and_spring_term(fix):
    """k < springsteps: DEV- carries 0.5 x the second difference (interior), W moves by
    0.05 DEV, errs is over 2M, x is from W- only, nothing is projected or redistributed."""
    bl = Blend(**{k: CANON[k] for k in ("N", "f", "alpha", "phidb", "chi", "ens")})
    g, st = _state(CANON, A0, B0, fix, springsteps=3)
    rep = st.iterate(0)
    a, b, want = string_iteration(bl, A0, B0, 0, springsteps=3, fix=fix, M=g.M)
    ga, gb = _rows(st, g.M)
    np.testing.assert_allclose(ga, a, rtol=1e-12, atol=1e-13)
    np.testing.assert_allclose(gb, b, rtol=1e-12, atol=1e-13)
    _check_report(rep, want, ("errs", "errs3", "x", "xref"))
    np.testing.assert_array_equal(rep["steps"], want["steps"])
    assert abs(rep["fitup"] - want["fitup"]) < 1e-13


@pytest.mark.parametrize("fix", [(1, 1), (0, 0)])
def test_collinear_string_projection_and_redistribution(fix):
    """k >= springsteps on a monotone uniform string: the tangent is parallel to DEV-, so the
    interior replicas keep their W- through the mixing and then move 0.05 of the way to the
    straight line; free ends move by 0.05 DEV- and are not redistributed."""
    bl = Blend(**{k: CANON[k] for k in ("N", "f", "alpha", "phidb", "chi", "ens")})
    g, st = _state(CANON, A0, B0, fix)
    rep = st.iterate(5)
    a, b, want = string_iteration(bl, A0, B0, 5, fix=fix, M=g.M)
    ga, gb = _rows(st, g.M)
    np.testing.assert_allclose(ga, a, rtol=1e-12, atol=1e-13)
    np.testing.assert_allclose(gb, b, rtol=1e-12, atol=1e-13)
    _check_report(rep, want, ("errs", "errs3", "x", "xref", "normDER", "normDEV", "dotWs", "dotWs2"))
    np.testing.assert_array_equal(rep["steps"], want["steps"])
    assert abs(rep["fitup"] - want["fitup"]) < 1e-13
    # the interior targets are exactly the straight line: a string already equally spaced
    # in W- stays put
    if fix == (1, 1):
        line = np.linspace(A0[0], A0[-1], R)
        g2, st2 = _state(CANON, line, B0, fix)
        st2.iterate(5)
        np.testing.assert_allclose(_rows(st2, g2.M)[0], line, rtol=0, atol=1e-13)


def test_grand_canonical_iteration_with_w_plus_relaxation():
    """ens 2: the uniform W+ relaxes over solve_field0's 19 sub-iterations with the adaptive
    lambda of scft.cu:516-554, so DEV+ stays non-zero and enters errs (over 2M) and errs3."""
    bl = Blend(**{k: GRAND[k] for k in ("N", "f", "alpha", "phidb", "chi", "ens")})
    b0 = np.zeros(R)
    g, st = _state(GRAND, A0, b0, (0, 0))
    rep = st.iterate(5)
    a, b, want = string_iteration(bl, A0, b0, 5, fix=(0, 0), M=g.M)
    assert np.all(np.abs(b - b0) > 1e-3)            # W+ did move
    ga, gb = _rows(st, g.M)
    np.testing.assert_allclose(ga, a, rtol=1e-11, atol=1e-12)
    np.testing.assert_allclose(gb, b, rtol=1e-11, atol=1e-12)
    _check_report(rep, want, ("errs", "errs3", "x", "dotWs", "normDEV"))
    np.testing.assert_array_equal(rep["steps"], want["steps"])


def test_arc_length_ignores_w_plus():
    """The string metric is W- only: replicas with different W+ give the same x."""
    _, st1 = _state(CANON, A0, B0, (1, 1))
    _, st2 = _state(CANON, A0, B0[::-1] * 3.0, (1, 1))
    r1, r2 = st1.iterate(5), st2.iterate(5)
    np.testing.assert_array_equal(r1["x"], r2["x"])


@pytest.mark.parametrize("run", [CANON, GRAND], ids=["canonical", "grand"])
def test_free_energy_uniform_string(run):
    """FreeE on a uniform string: per replica -lnQ + chi phiA phiB - wA phiA - wB phiB after
    solve_field0(100, 1e-4), phctot = phiA / f, and alf the cumulative W- distance."""
    bl = Blend(**{k: run[k] for k in ("N", "f", "alpha", "phidb", "chi", "ens")})
    g, st = _state(run, A0, B0, (1, 1))
    fe = st.free_energy()
    FE, FEs, ph, alf, b = free_energy(bl, A0, B0, M=g.M)
    np.testing.assert_allclose(fe["FEs"], FEs, rtol=1e-11, atol=1e-12)
    np.testing.assert_allclose(fe["phctot"], ph, rtol=1e-12)
    np.testing.assert_allclose(fe["alf"], alf, rtol=1e-12)
    assert abs(fe["FE"] - FE) < 1e-10
    np.testing.assert_allclose(_rows(st, g.M)[1], b, rtol=1e-11, atol=1e-12)


def test_main_replay_of_uniform_string_matches_oracle():
    """The closed-form replay of main() (prerelaxation, 52 iterations with FreeE at k = 0, 50
    and 51, final FreeE) against the oracle running the same sequence: the test-side driver
    replay is itself pinned. Grand canonical, so FreeE's W+ relaxation changes the later
    iterations and the cadence shows in the fields (the fixed ends' W+ moves only there)."""
    run = GRAND
    bl = Blend(**{k: run[k] for k in ("N", "f", "alpha", "phidb", "chi", "ens")})
    b0 = np.zeros(R)
    g, st = _state(run, A0, b0, (1, 1))
    for ii in range(R):
        g.solve_field0(st.W[ii], st.DEV[ii], run["chi"], run["f"], run["N"], run["alpha"], run["phidb"],
                       run["ens"], 200, 1e-4)
    maxiter = 52
    for k in range(maxiter):
        st.iterate(k, 1e-3)
        if k % 50 == 0 or k == maxiter - 1:
            st.free_energy()
    fe = st.free_energy()
    want = replay_main(bl, A0, b0, maxiter, 200, fix=(1, 1), M=g.M)
    assert want["FE_calls"] == [0, 50, 51]
    ga, gb = _rows(st, g.M)
    np.testing.assert_allclose(ga, want["a"], rtol=1e-10, atol=1e-11)
    np.testing.assert_allclose(gb, want["b"], rtol=1e-10, atol=1e-11)
    np.testing.assert_allclose(fe["FEs"], want["final"][1], rtol=1e-10, atol=1e-11)
    # without the FreeE at k = 50 the fields differ visibly: the cadence is observable
    alt = replay_main(bl, A0, b0, maxiter, 200, fix=(1, 1), M=g.M, skip=1000)
    assert np.max(np.abs(alt["b"] - want["b"])) > 1e-8
