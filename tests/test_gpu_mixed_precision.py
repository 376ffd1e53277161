"""Mixed precision: an FP32 context evaluates FreeE's free energy from an FP64 props on its
FP32-relaxed fields, so its F agrees with an FP64 context's to second order in the FP32
field error (bar: |dF| <= 1e-6 per replica and between replicas).

That argument needs F to be stationary in W+ at the W+ saddle. It holds in the canonical
ensemble when n1, n2 and Nh are even, because the Simpson weights then integrate exactly.
With an odd count the reference's weights make the copolymer normalisation differ from
phidb (density.h:405-422), so F drifts with the uniform part of W+, which canonical
relaxation never removes. In the grand-canonical ensemble the reference's F uses
Q2 + phidb*QC for lnQ (density.h:398), which is stationary only up to that same
discretisation; and there the densities are proportional to QC and Q2 themselves (zC = phidb,
zB = 1), so the FP32 error of the partition functions (~3e-7 relative) moves the FP32-relaxed
W+ and F at first order: its bar is 5e-6.

Without the FP64 props (fe_fp64 = 0, all-FP32) the bar is 1e-6 on the benchmark string and at
converged saddles: the FP32 propagator keeps the mean of q (the 3-D DC coefficient) in FP64
through each column pass, and the densities' Simpson sums are scaled to their exact means."""
import sys

import numpy as np
import pytest

from scft_string_amd import Context

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _string(m, D, R, seed=3):
    M = int(np.prod(m))
    rng = np.random.default_rng(seed)
    x = (np.arange(m[0]) * D[0] / m[0])[:, None, None] * np.ones(m)
    y = (np.arange(m[1]) * D[1] / m[1])[None, :, None] * np.ones(m)
    lam = 0.5 * np.cos(2 * np.pi * x / D[0])
    other = 0.5 * np.cos(2 * np.pi * x / D[0]) * np.cos(2 * np.pi * y / D[1])
    W = np.zeros((R, 2, M))
    for i in range(R):
        t = i / (R - 1)
        W[i, 0] = ((1 - t) * lam + t * other).reshape(-1) + 0.01 * rng.standard_normal(M)
    return W


@pytest.mark.parametrize("ens,f,bar", [(1, 0.5, 1e-6), (2, 0.45, 5e-6)])
def test_free_energy_fp32_context_matches_fp64(ens, f, bar):
    m, D, N, R = (32, 32, 32), (4.0, 4.0, 4.0), 64, 4
    phidb = 0.85 if ens == 1 else 0.5
    W = _string(m, D, R)
    fe = {}
    first_order = {}
    # evolve the string a little in FP32, then relax W+ of every replica (ends included)
    # tightly, so FreeE's own W+ sub-iterations start at the W+ saddle in both precisions
    with Context(m, D, N, f, 0.125, ens, phidb, 16.0, nrep=R, precision="fp32", fix=(1, 1)) as ctx:
        ctx.set_fields(W)
        for k in range(2):
            ctx.string_iterate(k)
        Wk = ctx.get_fields()
    with Context(m, D, N, f, 0.125, ens, phidb, 16.0, nrep=R, precision="fp64", fix=(1, 1)) as ctx:
        ctx.set_fields(Wk)
        ctx.solve_w_plus(3000, 1e-7)
        Wk = ctx.get_fields()
    # F of the same fields in both precisions
    for prec in ("fp32", "fp64"):
        with Context(m, D, N, f, 0.125, ens, phidb, 16.0, nrep=R, precision=prec, fix=(1, 1)) as ctx:
            ctx.set_fields(Wk)
            fe[prec] = ctx.free_energy()["FEs"]
        with Context(m, D, N, f, 0.125, ens, phidb, 16.0, nrep=R, precision=prec, fix=(1, 1),
                     fe_fp64=0) as ctx:
            ctx.set_fields(Wk)
            first_order[prec] = ctx.free_energy()["FEs"]
    d = np.abs(fe["fp32"] - fe["fp64"])
    between = np.abs((fe["fp32"] - fe["fp32"][0]) - (fe["fp64"] - fe["fp64"][0]))
    d1 = np.abs(first_order["fp32"] - first_order["fp64"])
    print(f"ens={ens}: FreeE |dF| max {d.max():.2e}, between replicas {between.max():.2e}; "
          f"all-FP32 FreeE |dF| max {d1.max():.2e}", flush=True)
    assert d.max() <= bar
    assert between.max() <= bar


def test_full_size_string_free_energy_fp32_within_1e6():
    """The benchmark string (16 replicas, 64^3, N = 128, chiN = 18, f = 0.45, phidb = 0.9,
    canonical, fix = (1, 1), lamellar -> cylinder ends), 3 iterations with the FP32
    propagator; then FreeE on those identical fields in the FP32 context (W+ relaxed with FP32
    props, F from an FP64 props: fe_fp64) and in an FP64 context. |dF| <= 1e-6 per replica and
    between replicas, and also for the all-FP32 FreeE (fe_fp64 = 0)."""
    from bench import CFG, initial_fields

    c = CFG
    R = c["nrep"]
    with Context(c["m"], c["D"], c["N"], c["f"], c["alpha"], c["ens"], c["phidb"], c["chi"], nrep=R,
                 precision="fp32", fix=c["fix"]) as ctx:
        ctx.set_fields(initial_fields(c["m"], c["D"], R))
        for k in range(3):
            ctx.string_iterate(k)
        W = ctx.get_fields()
    fe = {}
    for prec, fe64 in (("fp32", 1), ("fp64", 1), ("fp32", 0)):
        with Context(c["m"], c["D"], c["N"], c["f"], c["alpha"], c["ens"], c["phidb"], c["chi"], nrep=R,
                     precision=prec, fix=c["fix"], fe_fp64=fe64) as ctx:
            ctx.set_fields(W)
            fe[(prec, fe64)] = ctx.free_energy()["FEs"]
    d = np.abs(fe[("fp32", 1)] - fe[("fp64", 1)])
    rel = (fe[("fp32", 1)] - fe[("fp32", 1)][0]) - (fe[("fp64", 1)] - fe[("fp64", 1)][0])
    d32 = np.abs(fe[("fp32", 0)] - fe[("fp64", 1)])
    _say(f"16 x 64^3 string: |dF| {d.max():.2e}, between replicas {np.abs(rel).max():.2e}; "
         f"all-FP32 FreeE |dF| {d32.max():.2e}")
    assert d.max() <= 1e-6 and np.abs(rel).max() <= 1e-6
    # all-FP32: the lnQ error of the FP32 propagator (measured 3.7e-7; 2.9e-6 before the DC
    # coefficient was carried in FP64, 2.5e-5 before the FP32 Simpson sums were scaled to
    # their exact means)
    assert d32.max() <= 1e-6


def test_fp32_props_free_energy_within_1e6_on_the_benchmark_string():
    """One props of all 16 replicas of the benchmark string's starting fields (16 x 64^3,
    N = 128) in an FP32 and an FP64 context, and F of each from its own densities and lnQ
    (scft.cu:812-853 without the W+ relaxation): |dF| <= 1e-6 for every replica. This is the
    all-FP32 propagator error itself (measured 3.2e-7; 2.4e-6 with the 3-D DC coefficient
    rounded to FP32 in every column pass)."""
    from bench import CFG, free_energies, initial_fields

    c = CFG
    W = initial_fields(c["m"], c["D"], 16)
    F = {}
    for prec in ("fp32", "fp64"):
        with Context(c["m"], c["D"], c["N"], c["f"], c["alpha"], c["ens"], c["phidb"], c["chi"], nrep=16,
                     precision=prec, fix=(0, 0)) as ctx:
            phi, lnQ, _ = ctx.props_batch(W)
        F[prec] = free_energies(W, phi, lnQ, c["chi"])
    d = np.abs(F["fp32"] - F["fp64"])
    _say(f"16 x 64^3 string, one props: all-FP32 |dF| max {d.max():.2e}, rms {np.sqrt(np.mean(d ** 2)):.2e}")
    assert d.max() <= 1e-6


@pytest.mark.parametrize("N,f", [(64, 0.5), (128, 0.4)])
def test_fp32_density_normalisation_matches_fp64(N, f):
    """FP32 contexts scale each Simpson sum (copolymer A and B blocks, B homopolymer) to its
    exact mean, which the symmetric-step identity <q1'[s] q2'[N-s]> = QC' for every s fixes:
    the copolymer and homopolymer volume fractions then equal the FP64 context's to FP64
    rounding, the odd-count weight quirk of N = 128, f = 0.4 (n1 = 51) included."""
    m, D, R = (32, 32, 32), (4.0, 4.0, 4.0), 3
    W = _string(m, D, R)
    means = {}
    for prec in ("fp32", "fp64"):
        with Context(m, D, N, f, 0.125, 1, 0.85, 16.0, nrep=R, precision=prec) as ctx:
            phi, lnQ, phis = ctx.props_batch(W)
        means[prec] = np.stack([phi[:, 0].mean(1), phi[:, 3].mean(1), phi[:, 2].mean(1)], axis=1)
        means[prec + "_phis"] = phis
    d = np.max(np.abs(means["fp32"] - means["fp64"]))
    dp = np.max(np.abs(means["fp32_phis"] - means["fp64_phis"]))
    _say(f"N={N} f={f}: FP32 vs FP64 block means {d:.1e}, phis {dp:.1e}")
    assert d <= 1e-12 and dp <= 1e-12


def _lamellae(m, D):
    """x-, y- and z-lamellae of one period: three saddles of equal F in a cubic box"""
    M = int(np.prod(m))
    W = np.zeros((3, 2, M))
    for i in range(3):
        ax = np.arange(m[i]) * D[i] / m[i]
        shape = [1, 1, 1]
        shape[i] = m[i]
        W[i, 0] = (1.5 * np.cos(2 * np.pi * ax / D[i])).reshape(shape).repeat(m[0] // shape[0], 0) \
            .repeat(m[1] // shape[1], 1).repeat(m[2] // shape[2], 2).reshape(-1)
    return W


def test_saddles_converged_independently_in_fp32_and_fp64():
    """Three saddle points (x-, y- and z-lamellae in a 32^3 box of side 1.5, one lamellar
    period: chiN = 15, f = 0.5, phidb = 0.8, N = 40, canonical; a box of side 4.5 relaxes a
    one-period cosine to the disordered state) converged separately in FP32 and in FP64 from the
    same start, as independent replicas with Anderson mixing (H = 20) until errs stops
    falling. FreeE of each: with F from an FP64 props on the FP32-converged fields (fe_fp64)
    |dF| <= 1e-6 per replica and between replicas, since F is stationary at the saddle and
    the FP32 field error enters at second order; the all-FP32 F carries the FP32 error of lnQ
    at first order and is bounded by 1e-6 too (measured 6.6e-7)."""
    m, D, N = (32, 32, 32), (1.5, 1.5, 1.5), 40
    args = (m, D, N, 0.5, 0.1, 1, 0.8, 15.0)
    W = _lamellae(m, D)
    fe, errs = {}, {}
    for prec in ("fp32", "fp64"):
        with Context(*args, nrep=3, precision=prec, dostring=0, anderson=20) as ctx:
            ctx.set_fields(W)
            best, since = np.inf, 0
            for k in range(3000):
                e = float(np.max(ctx.string_iterate(k, 1e-6)["errs"]))
                if e < 0.9 * best:
                    best, since = e, 0
                else:
                    since += 1
                if e < 1e-9 or since > 100:
                    break
            errs[prec] = (best, k + 1)
            Wc = ctx.get_fields()
            fe[prec] = ctx.free_energy()["FEs"]
            amp = np.ptp(Wc.reshape(3, 2, -1)[:, 0], axis=1)
        if prec == "fp32":
            with Context(*args, nrep=3, precision=prec, dostring=0, fe_fp64=0) as ctx:
                ctx.set_fields(Wc)
                fe["fp32_all"] = ctx.free_energy()["FEs"]
    d = fe["fp32"] - fe["fp64"]
    d_all = fe["fp32_all"] - fe["fp64"]
    _say(f"errs fp32 {errs['fp32']}, fp64 {errs['fp64']}; FEs fp64 {fe['fp64']}; |dF| {np.abs(d).max():.1e} "
         f"between {np.abs(d - d[0]).max():.1e}; all-FP32 |dF| {np.abs(d_all).max():.1e} "
         f"between {np.abs(d_all - d_all[0]).max():.1e}")
    assert errs["fp64"][0] < 1e-8 and errs["fp32"][0] < 1e-4
    assert np.ptp(fe["fp64"]) < 1e-9                      # three symmetric saddles
    assert np.all(fe["fp64"] < 3.55) and np.all(amp > 5.0)  # ordered (disordered: F = 3.6, flat W-)
    assert np.abs(d).max() <= 1e-6 and np.abs(d - d[0]).max() <= 1e-6
    assert np.abs(d_all).max() <= 1e-6 and np.abs(d_all - d_all[0]).max() <= 1e-6


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_fp32_props_of_a_lamella_along_each_axis_within_1e6(axis):
    """A lamella along one axis makes every line (x) or every plane (y, z) of a pass carry the
    same data, so any rounding of the mean that a pass makes is the same everywhere: the
    coherent case. One all-FP32 props at the benchmark's grid and contour length (64^3,
    N = 128) against FP64: |dF| <= 1e-6. lnQ carries more of the FP32 error than F does
    (measured |d lnQ| 1.5e-6 along x with the two-pass step, 1.7e-6 with the four-launch
    step, against |dF| 2.6e-7 and 3.4e-7): its bar is 3e-6."""
    m, D, N = (64, 64, 64), (6.0, 6.0, 6.0), 128
    M = int(np.prod(m))
    g = np.meshgrid(*[2 * np.pi * np.arange(m[i]) / m[i] for i in range(3)], indexing="ij")
    W = np.zeros((1, 2, M))
    W[0, 0] = (2.0 * np.cos(g[axis])).reshape(-1)
    res = {}
    for prec in ("fp32", "fp64"):
        with Context(m, D, N, 0.45, 0.1, 1, 0.9, 18.0, nrep=1, precision=prec) as ctx:
            phi, lnQ, _ = ctx.props_batch(W)
        wa, wb = W[0, 0] + W[0, 1], W[0, 1] - W[0, 0]
        F = -lnQ[0] + np.mean(18.0 * phi[0, 0] * phi[0, 1] - wa * phi[0, 0] - wb * phi[0, 1])
        res[prec] = (lnQ[0], F)
    dF = abs(res["fp32"][1] - res["fp64"][1])
    dl = abs(res["fp32"][0] - res["fp64"][0])
    _say(f"lamella along axis {axis}: all-FP32 |dF| {dF:.2e}, |dlnQ| {dl:.2e} (lnQ {res['fp64'][0]:.4f})")
    assert dF <= 1e-6 and dl <= 3e-6


def test_string_converged_independently_in_fp32_and_fp64():
    """A 5-replica string between an x- and a y-lamella (32^3 box of side 1.5, one lamellar
    period; chiN = 15, f = 0.5, phidb = 0.8, N = 40, canonical), both ends fixed at the
    FP64-converged lamellae, converged separately in FP32 and in FP64 from the same linear
    interpolation for 1500 iterations: FP64 errs3 falls to ~1e-14, FP32 to its rounding floor
    (~1e-6). FreeE of the two strings within 1e-6 of each other per replica and between
    replicas: the FP32 context with its final props in FP64 (the default; measured 1.4e-7)
    and all-FP32 (fe_fp64 = 0; 8.4e-7, its FP32 lnQ error entering at first order). On a 16^3
    grid, which runs the small-grid LDS kernels without FP64 line sums, all-FP32 gives 1.2e-6."""
    m, Dl, N = (32, 32, 32), 1.5, 40
    D = (Dl, Dl, Dl)
    args = (m, D, N, 0.5, 0.1, 1, 0.8, 15.0)
    M = int(np.prod(m))
    x, y, _ = np.meshgrid(*[np.arange(m[i]) * D[i] / m[i] for i in range(3)], indexing="ij")
    W = np.zeros((2, 2, M))
    W[0, 0] = 2.0 * np.cos(2 * np.pi * x / Dl).reshape(-1)
    W[1, 0] = 2.0 * np.cos(2 * np.pi * y / Dl).reshape(-1)
    with Context(*args, nrep=2, precision="fp64", dostring=0, anderson=20) as ctx:
        ctx.set_fields(W)
        for k in range(1500):
            if float(np.max(ctx.string_iterate(k, 1e-10)["errs"])) < 1e-11:
                break
        Wend = ctx.get_fields().reshape(2, 2, M)
    R = 5
    W0 = np.stack([(1 - t) * Wend[0] + t * Wend[1] for t in np.linspace(0, 1, R)])
    fe, e3 = {}, {}
    for prec in ("fp64", "fp32"):
        with Context(*args, nrep=R, precision=prec, dostring=1, fix=(1, 1)) as ctx:
            ctx.set_fields(W0)
            for k in range(1500):
                rep = ctx.string_iterate(k, 1e-6)
            e3[prec] = float(np.max(rep["errs3"]))
            fe[prec] = ctx.free_energy()["FEs"]
            Wc = ctx.get_fields()
        if prec == "fp32":
            with Context(*args, nrep=R, precision=prec, dostring=1, fix=(1, 1), fe_fp64=0) as ctx:
                ctx.set_fields(Wc)
                fe["fp32_all"] = ctx.free_energy()["FEs"]
    d = fe["fp32"] - fe["fp64"]
    d_all = fe["fp32_all"] - fe["fp64"]
    _say(f"string errs3 fp64 {e3['fp64']:.1e} fp32 {e3['fp32']:.1e}; FEs fp64 {fe['fp64']}; |dF| {np.abs(d).max():.1e} "
         f"between {np.abs(d - d[0]).max():.1e}; all-FP32 |dF| {np.abs(d_all).max():.1e} "
         f"between {np.abs(d_all - d_all[0]).max():.1e}")
    assert e3["fp64"] < 1e-10 and e3["fp32"] < 1e-5
    assert fe["fp64"][0] < 3.55 and np.all(fe["fp64"][1:-1] > fe["fp64"][0] + 1e-3)  # ordered ends, a barrier
    assert np.abs(d).max() <= 1e-6 and np.abs(d - d[0]).max() <= 1e-6
    assert np.abs(d_all).max() <= 1e-6 and np.abs(d_all - d_all[0]).max() <= 1e-6
