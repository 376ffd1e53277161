"""Closed forms of the string method for spatially uniform fields.

When every replica's W- and W+ are constant in space, each Richardson step multiplies a
constant propagator by exactly exp(-w ds) (both Richardson terms carry E1/4 E1/2 E1/4 =
E1/2 E1/2 = e^{-w ds}; the kinetic factor is 1 on the k = 0 mode), so

    q1[s] q2[N-s] = QC = exp(-(n1 wA + n2 wB)/N)   for every s,     Q2 = exp(-Nh wB / N),

with wA = W- + W+, wB = W+ - W- (scft.cu:823). The Simpson sums of density.h:401-430 then
reduce to weight totals, and every quantity of solve_field0 (scft.cu:516-554), one
solve_field iteration (scft.cu:592-757) and FreeE (scft.cu:812-853) is a scalar per
replica. The spline of a monotone uniform string is exactly linear in the arc-length
coordinate, because x_i = (a_i - a_0)/(a_{R-1} - a_0) when the metric is |W-| distance.

These formulas are the mathematics of the uniform case, derived here from the densities'
definition; the tests compare the oracle and the GPU library with them. They are not a
transcription of either implementation: no FFT, no contour loop, no spline code.
"""
import math

import numpy as np


def simpson_weight_total(n):
    """2 + sum_{s=1}^{n-1} w_s with w = 4, 2, 4, ... starting at 4 (density.h:405-422)."""
    return 2 + sum(4 if s % 2 == 1 else 2 for s in range(1, n))


class Blend:
    """Run constants of the AB diblock + B homopolymer blend (density.h:344-399)."""

    def __init__(self, N, f, alpha, phidb, chi, ens):
        self.N, self.alpha, self.phidb, self.chi, self.ens = N, alpha, phidb, chi, ens
        self.n1 = int(round(N * f))            # density.h:346
        self.n2 = N - self.n1
        self.Nh = int(alpha * N)               # truncation, survey A.3
        self.f = self.n1 / N                   # scft.cu:1008

    def densities(self, a, b):
        """(phiA, phiB, lnQ, QC, Q2) of a uniform field W- = a, W+ = b."""
        N = self.N
        wA, wB = a + b, b - a
        QC = math.exp(-(self.n1 * wA + self.n2 * wB) / N)
        Q2 = math.exp(-self.Nh * wB / N)
        sC1, sC2, sH = (simpson_weight_total(n) / (3.0 * N) for n in (self.n1, self.n2, self.Nh))
        if self.ens == 1:
            zC, zB = self.phidb / QC, (1 - self.phidb) / (self.alpha * Q2)
            lnQ = self.phidb * math.log(QC) + (1 - self.phidb) / self.alpha * math.log(Q2)
        else:
            zC, zB = self.phidb, 1.0
            lnQ = Q2 + self.phidb * QC        # density.h:398, not a logarithm
        phiA = zC * QC * sC1
        phiB = zC * QC * sC2 + zB * Q2 * sH
        return phiA, phiB, lnQ, QC, Q2

    def solve_w_plus(self, a, b, max_iter, err_tol):
        """solve_field0 (scft.cu:516-554) on a uniform replica: returns (k, b)."""
        lam, errt, err, k = 0.05, 1.0, 1.0, 1
        while k < max_iter and err > err_tol:
            pA, pB, *_ = self.densities(a, b)
            d = self.chi * (pA + pB - 1.0)
            err = abs(d)                       # sqrt(sum_r d^2 / M)
            b += lam * d
            lam = min(0.1, lam * 1.09) if err < errt else max(0.005, lam / 1.10)
            errt = err
            k += 1
        return k, b


def arc_positions(a):
    """Normalised cumulative |W-| distance (scft.cu:612-622) of a uniform string."""
    d = np.abs(np.diff(a))
    x = np.zeros(len(a))
    x[1:-1] = np.cumsum(d / d.sum())[:-1]
    x[-1] = 1.0
    return x


def string_iteration(bl, a, b, k, springsteps=0, fix=(1, 1), M=1):
    """One solve_field iteration (scft.cu:592-757) of a uniform string.

    a, b: W- and W+ per replica (monotone a). Returns the new (a, b) and a report with the
    quantities scft_iter_report carries. Derivation, replica by replica:
      DEV- = chi (phiB - phiA) - 2a (+ 0.5 (a_{i+1} + a_{i-1} - 2 a_i) in the spring phase),
      DEV+ = chi (phiA + phiB - 1) after solve_field0(20, 1e-3);
      errs = sqrt((DEV-^2 + DEV+^2) / 2);
      in the projection phase the tangent dW-/dx is a_{R-1} - a_0 everywhere (the spline of
      linear data), DEV- is parallel to it, so the projection removes all of DEV- for
      interior replicas: dotWs = sign(DEV- * (a_{R-1} - a_0)), errs3 = |DEV+| / sqrt(2);
      W += 0.05 DEV; x from the new a; then interior replicas move 0.05 of the way to the
      straight line through the new ends at x = i/(R-1).
    """
    R = len(a)
    a = np.asarray(a, dtype=float).copy()
    b = np.asarray(b, dtype=float).copy()
    act = np.ones(R, dtype=bool)
    if fix[0]:
        act[0] = False
    if fix[1]:
        act[-1] = False
    spring = k < springsteps
    dm, dp, steps = np.zeros(R), np.zeros(R), np.zeros(R, dtype=int)
    for i in range(R):
        if not act[i]:
            continue
        kk, b[i] = bl.solve_w_plus(a[i], b[i], 20, 1e-3)
        steps[i] = 1 + kk
        pA, pB, *_ = bl.densities(a[i], b[i])
        dm[i] = bl.chi * (pB - pA) - 2.0 * a[i]
        if spring and 0 < i < R - 1:
            dm[i] += 0.5 * (a[i + 1] + a[i - 1] - 2.0 * a[i])
        dp[i] = bl.chi * (pA + pB - 1.0)
    errs = np.sqrt((dm ** 2 + dp ** 2) / 2.0)
    rep = {"errs": errs, "steps": steps}
    slope = a[-1] - a[0]
    if not spring:
        rep["normDER"] = np.where(act, math.sqrt(M) * abs(slope), 0.0)
        rep["normDEV"] = np.where(act, math.sqrt(M) * np.abs(dm), 0.0)
        rep["dotWs"] = np.where(act, np.sign(dm * slope), 0.0)
        inner = act.copy()
        inner[0] = inner[-1] = False
        dm = np.where(inner, 0.0, dm)
        rep["dotWs2"] = np.where(inner, 0.0, rep["dotWs"])
        errs3 = np.where(inner, np.abs(dp) / math.sqrt(2.0), 0.0)
    else:
        errs3 = errs.copy()
    errs3[0], errs3[-1] = errs[0], errs[-1]
    rep["errs3"] = errs3
    a = a + 0.05 * dm
    b = b + 0.05 * dp
    x = arc_positions(a)
    xref = np.arange(R) / (R - 1.0)
    rep["x"], rep["xref"] = x, xref
    rep["fitup"] = math.sqrt(np.sum((x[1:-1] - xref[1:-1]) ** 2) / (R - 2.0)) / (R - 1.0)
    if not spring:
        line = a[0] + (a[-1] - a[0]) * xref
        for i in range(1, R - 1):
            if act[i]:
                a[i] += (line[i] - a[i]) * 0.05
    return a, b, rep


def free_energy(bl, a, b, M=1):
    """FreeE (scft.cu:812-853) of a uniform string, every replica (single process).

    Returns (FE, FEs, phctot, alf, b) with alf the cumulative |W-| distance (not normalised)
    and b after solve_field0(100, 1e-4)."""
    R = len(a)
    b = np.asarray(b, dtype=float).copy()
    FEs, ph = np.zeros(R), np.zeros(R)
    for i in range(R):
        _, b[i] = bl.solve_w_plus(a[i], b[i], 100, 1e-4)
        pA, pB, lnQ, *_ = bl.densities(a[i], b[i])
        wA, wB = a[i] + b[i], b[i] - a[i]
        FEs[i] = -lnQ + bl.chi * pA * pB - wA * pA - wB * pB
        ph[i] = pA / bl.f
    alf = np.concatenate([[0.0], np.cumsum(math.sqrt(M) * np.abs(np.diff(a)))])
    return FEs.sum(), FEs, ph, alf, b


def replay_main(bl, a, b, maxiter, prerelax, fix=(0, 0), M=1, skip=50):
    """main()'s order (scft.cu:1102-1132) on a uniform string: solve_field0(prerelax, 1e-4) on
    every replica, maxiter solve_field iterations with the periodic block (scft.cu:760-791) at
    k % 50 == 0 and k == maxiter - 1 (the check values first, with the FE of the previous
    FreeE, then FreeE, which relaxes W+ of every replica), then the final FreeE.

    Returns a dict: check = (k, errtot, FE, errs, errs3) of the last periodic block, FE_calls =
    the iterations FreeE ran at, final = (FE, FEs, phctot, alf) and the final a, b."""
    a = np.asarray(a, dtype=float).copy()
    b = np.asarray(b, dtype=float).copy()
    for i in range(len(a)):
        _, b[i] = bl.solve_w_plus(a[i], b[i], prerelax, 1e-4)
    FE, check, calls = 0.0, None, []
    for k in range(maxiter):
        a, b, rep = string_iteration(bl, a, b, k, fix=fix, M=M)
        if k % skip == 0 or k == maxiter - 1:
            errtot = math.sqrt(np.sum(rep["errs3"] ** 2) / len(a))
            check = (k, errtot, FE, rep["errs"].copy(), rep["errs3"].copy())
            FE, _, _, _, b = free_energy(bl, a, b, M=M)
            calls.append(k)
    FE, FEs, ph, alf, b = free_energy(bl, a, b, M=M)
    return {"check": check, "FE_calls": calls, "final": (FE, FEs, ph, alf), "a": a, "b": b}
