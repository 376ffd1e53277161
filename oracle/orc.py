"""ctypes binding of the FP64 CPU oracle (oracle/liborc.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker. The product package never imports it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

NREP_MAX = 256


class OrcRun(C.Structure):
    _fields_ = [
        ("chi", C.c_double), ("f", C.c_double), ("alpha", C.c_double), ("phidb", C.c_double),
        ("N", C.c_int), ("ens", C.c_int), ("nrep", C.c_int), ("dostring", C.c_int),
        ("fix", C.c_int * 2), ("springsteps", C.c_int),
    ]


class OrcIterReport(C.Structure):
    _fields_ = [
        ("errs", C.c_double * NREP_MAX), ("errs3", C.c_double * NREP_MAX),
        ("x", C.c_double * NREP_MAX), ("xref", C.c_double * NREP_MAX),
        ("dotWs", C.c_double * NREP_MAX), ("dotWs2", C.c_double * NREP_MAX),
        ("normDER", C.c_double * NREP_MAX), ("normDEV", C.c_double * NREP_MAX),
        ("steps", C.c_int * NREP_MAX), ("fitup", C.c_double),
    ]


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        # ORC_LIB names another build of the oracle: the mutation tests point it at copies
        # built with ORC_MUTANT set, each planting one known mistake.
        path = os.environ.get("ORC_LIB") or os.path.join(_HERE, "liborc.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        dp = C.POINTER(C.c_double)
        ip = C.POINTER(C.c_int)
        vp = C.c_void_p
        L.orc_grid_create.restype = vp
        L.orc_grid_create.argtypes = [ip, dp]
        L.orc_grid_destroy.argtypes = [vp]
        L.orc_grid_M.argtypes = [vp]
        L.orc_grid_Mk.argtypes = [vp]
        L.orc_set_threads.argtypes = [C.c_int]
        L.orc_ksq.argtypes = [vp, dp]
        L.orc_rfft3.argtypes = [vp, dp, dp]
        L.orc_irfft3.argtypes = [vp, dp, dp]
        L.orc_step.argtypes = [vp, dp, dp, dp, dp, dp, dp]
        L.orc_props.argtypes = [vp, dp, dp, dp, C.c_int, C.c_double, C.c_double, C.c_double,
                                C.c_int, dp, dp]
        L.orc_solve_field0.restype = C.c_int
        L.orc_solve_field0.argtypes = [vp, dp, dp, C.c_double, C.c_double, C.c_int, C.c_double,
                                       C.c_double, C.c_int, C.c_int, C.c_double]
        L.orc_spline_coeffs.argtypes = [C.c_int, dp, dp, dp, dp, dp]
        L.orc_spline_deriv_nodes.argtypes = [C.c_int, dp, dp, dp]
        L.orc_spline_value_at.argtypes = [C.c_int, dp, dp, dp, dp]
        L.orc_edist.restype = C.c_double
        L.orc_edist.argtypes = [C.c_int, dp, dp]
        L.orc_string_iter.argtypes = [vp, C.POINTER(OrcRun), C.POINTER(dp), C.POINTER(dp), ip,
                                      C.c_int, C.c_double, dp, C.POINTER(OrcIterReport)]
        L.orc_free_energy.restype = C.c_double
        L.orc_free_energy.argtypes = [vp, C.POINTER(OrcRun), C.POINTER(dp), C.POINTER(dp), ip, dp,
                                      dp, dp, dp, dp]
        L.orc_dlnq_dbox.argtypes = [vp, dp, C.c_int, C.c_double, C.c_double, C.c_double, C.c_int,
                                    dp, dp]
        L.orc_grid_set_phiha.argtypes = [vp, C.c_double]
        L.orc_anderson_create.restype = vp
        L.orc_anderson_create.argtypes = [C.c_int, C.c_int]
        L.orc_anderson_destroy.argtypes = [vp]
        L.orc_anderson_iter.restype = C.c_double
        L.orc_anderson_iter.argtypes = [vp, dp, vp, C.c_double, C.c_double, C.c_int, C.c_double, C.c_double,
                                        C.c_int, C.c_double, C.c_double]
        L.orc_small_solve.argtypes = [C.c_int, dp, dp, dp]
        _LIB = L
    return _LIB


def _dp(a):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.POINTER(C.c_double))


def set_threads(n):
    lib().orc_set_threads(int(n))


class Grid:
    """One oracle grid (m, D) with its FFT plans and work buffers."""

    def __init__(self, m, D):
        self.m = tuple(int(v) for v in m)
        self.D = tuple(float(v) for v in D)
        mm = (C.c_int * 3)(*self.m)
        dd = (C.c_double * 3)(*self.D)
        self._h = lib().orc_grid_create(mm, dd)
        self.M = lib().orc_grid_M(self._h)
        self.Mk = lib().orc_grid_Mk(self._h)

    def close(self):
        if self._h:
            lib().orc_grid_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def ksq(self):
        out = np.empty(self.Mk)
        lib().orc_ksq(self._h, _dp(out))
        return out.reshape(self.m[0], self.m[1], self.m[2] // 2 + 1)

    def rfft3(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
        out = np.empty(2 * self.Mk)
        lib().orc_rfft3(self._h, _dp(a), _dp(out))
        return out.view(np.complex128).reshape(self.m[0], self.m[1], self.m[2] // 2 + 1)

    def irfft3(self, h):
        h = np.ascontiguousarray(h, dtype=np.complex128).reshape(-1)
        out = np.empty(self.M)
        lib().orc_irfft3(self._h, _dp(h.view(np.float64)), _dp(out))
        return out.reshape(self.m)

    def step(self, e4, e2, k12, k6, qin):
        out = np.empty(self.M)
        args = [np.ascontiguousarray(v, dtype=np.float64).reshape(-1) for v in (e4, e2, k12, k6, qin)]
        lib().orc_step(self._h, *[_dp(v) for v in args], _dp(out))
        return out.reshape(self.m)

    def props(self, W, N, alpha, phic, f, ens, want_q=False):
        """W: (2, M) W-, W+.  Returns dict with phi (4, M), lnQ, phis (5), Q (Q1,Q2,QC)."""
        M = self.M
        w = np.zeros(6 * M)
        w[: 2 * M] = np.asarray(W, dtype=np.float64).reshape(-1)
        lnQ = C.c_double()
        phis = np.zeros(6)
        Q = np.zeros(3)
        qs = np.empty(2 * (N + 1) * M) if want_q else None
        lib().orc_props(self._h, _dp(w), C.byref(lnQ), _dp(phis), int(N), float(alpha), float(phic),
                        float(f), int(ens), _dp(Q), _dp(qs) if want_q else None)
        out = {"phi": w[2 * M:].reshape(4, M).copy(), "lnQ": lnQ.value, "phis": phis[:5].copy(),
               "Q": Q.copy()}
        if want_q:
            out["q"] = qs.reshape(N + 1, 2, M)
        return out

    def set_phihA(self, phihA):
        """A homopolymer for later calls (orc_grid_set_phiha): ens 1 volume fraction, ens 2 activity."""
        lib().orc_grid_set_phiha(self._h, float(phihA))

    def dlnq_dbox(self, W, N, alpha, phic, f, ens):
        """Exact d lnQ / dD_i (i = x, y, z) of the discrete props at fixed W: (lnQ, (3,))."""
        M = self.M
        w = np.zeros(6 * M)
        w[: 2 * M] = np.asarray(W, dtype=np.float64).reshape(-1)
        lnQ = C.c_double()
        d = np.zeros(3)
        lib().orc_dlnq_dbox(self._h, _dp(w), int(N), float(alpha), float(phic), float(f), int(ens),
                            C.byref(lnQ), _dp(d))
        return lnQ.value, d

    def solve_field0(self, Wrow, DEV, chi, f, N, alpha, phic, ens, maxIter, errTol):
        return lib().orc_solve_field0(self._h, _dp(Wrow), _dp(DEV), float(chi), float(f), int(N),
                                      float(alpha), float(phic), int(ens), int(maxIter), float(errTol))


def spline_coeffs(x, a):
    n = len(x)
    x = np.ascontiguousarray(x, dtype=np.float64)
    a = np.ascontiguousarray(a, dtype=np.float64)
    c, b, d = np.zeros(n), np.zeros(n), np.zeros(n)
    lib().orc_spline_coeffs(n, _dp(x), _dp(a), _dp(c), _dp(b), _dp(d))
    return c, b, d


def spline_deriv_nodes(x, a):
    x = np.ascontiguousarray(x, dtype=np.float64)
    a = np.ascontiguousarray(a, dtype=np.float64)
    out = np.zeros(len(x))
    lib().orc_spline_deriv_nodes(len(x), _dp(x), _dp(a), _dp(out))
    return out


def spline_value_at(x, a, xr):
    x = np.ascontiguousarray(x, dtype=np.float64)
    a = np.ascontiguousarray(a, dtype=np.float64)
    xr = np.ascontiguousarray(xr, dtype=np.float64)
    out = np.zeros(len(x))
    lib().orc_spline_value_at(len(x), _dp(x), _dp(a), _dp(xr), _dp(out))
    return out


class StringState:
    """Host-side replica rows W[ii] (7M each) and DEV[ii] (2M each), as in scft.cu:1029-1051."""

    def __init__(self, grid, run_kw, W0):
        self.g = grid
        M = grid.M
        R = W0.shape[0]
        self.run = OrcRun(chi=run_kw["chi"], f=run_kw["f"], alpha=run_kw["alpha"],
                          phidb=run_kw["phidb"], N=run_kw["N"], ens=run_kw["ens"], nrep=R,
                          dostring=run_kw.get("dostring", 1),
                          fix=(C.c_int * 2)(*run_kw.get("fix", (0, 0))),
                          springsteps=run_kw.get("springsteps", 0))
        self.W = np.zeros((R, 7 * M))
        self.W[:, : 2 * M] = W0.reshape(R, 2 * M)
        self.DEV = np.zeros((R, 2 * M))
        self.doiis = np.ones(R, dtype=np.int32)
        fix = run_kw.get("fix", (0, 0))
        if fix[0] == 1:
            self.doiis[0] = 0
        if fix[1] == 1:
            self.doiis[R - 1] = 0
        self.alf = np.zeros(R)

    def _rows(self, arr):
        dp = C.POINTER(C.c_double)
        return (dp * arr.shape[0])(*[arr[i].ctypes.data_as(dp) for i in range(arr.shape[0])])

    def iterate(self, k, errTol=1e-3):
        rep = OrcIterReport()
        lib().orc_string_iter(self.g._h, C.byref(self.run), self._rows(self.W), self._rows(self.DEV),
                              self.doiis.ctypes.data_as(C.POINTER(C.c_int)), int(k), float(errTol),
                              _dp(self.alf), C.byref(rep))
        R = self.W.shape[0]
        return {key: np.array(getattr(rep, key)[:R]) for key in
                ("errs", "errs3", "x", "xref", "dotWs", "dotWs2", "normDER", "normDEV", "steps")} | {
                    "fitup": rep.fitup}

    def free_energy(self):
        R = self.W.shape[0]
        FEs0, FEs1, FEs, ph = np.zeros(R), np.zeros(R), np.zeros(R), np.zeros(R)
        FE = lib().orc_free_energy(self.g._h, C.byref(self.run), self._rows(self.W), self._rows(self.DEV),
                                   self.doiis.ctypes.data_as(C.POINTER(C.c_int)), _dp(self.alf),
                                   _dp(FEs0), _dp(FEs1), _dp(FEs), _dp(ph))
        return {"FE": FE, "FEs0": FEs0, "FEs1": FEs1, "FEs": FEs, "phctot": ph, "alf": self.alf.copy()}


class Anderson:
    """Anderson mixing state of one replica (orc_anderson_*). iterate() updates W (2, M) in place."""

    def __init__(self, grid, H, chi, f, N, alpha, phic, ens, lam, start):
        self.g = grid
        self.args = (float(chi), float(f), int(N), float(alpha), float(phic), int(ens), float(lam), float(start))
        self._h = lib().orc_anderson_create(grid.M, int(H))

    def iterate(self, W):
        M = self.g.M
        row = np.zeros(7 * M)
        row[: 2 * M] = W.reshape(-1)
        errs = lib().orc_anderson_iter(self.g._h, _dp(row), self._h, *self.args)
        W.reshape(-1)[:] = row[: 2 * M]
        return errs

    def close(self):
        if getattr(self, "_h", None):
            lib().orc_anderson_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()
