"""ctypes binding of libscft_hip.so (include/scft_api.h).

The product path is the HIP library; there is no CPU fallback. Loading raises if the
library has not been built (python -m scft_string_amd.build).
"""
import ctypes as C
import os

import numpy as np

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "libscft_hip.so")

(SCFT_OK, SCFT_ERR_INVALID, SCFT_ERR_UNSUPPORTED, SCFT_ERR_HIP, SCFT_ERR_NOMEM, SCFT_ERR_NAN,
 SCFT_ERR_NOCONV) = range(7)
SCFT_FP64, SCFT_FP32 = 0, 1

EXPORTED = [
    "scft_ctx_create", "scft_ctx_destroy", "scft_last_error", "scft_sync", "scft_props_batch",
    "scft_props_batch_device", "scft_set_fields", "scft_get_fields", "scft_get_densities",
    "scft_get_tangent", "scft_solve_w_plus", "scft_string_iterate", "scft_free_energy",
    "scft_debug_qq_invariant", "scft_debug_time_props", "scft_fft_path", "scft_dlnq_dbox",
    "scft_box_stress", "scft_set_box", "scft_relax_box", "scft_iterate_concurrent", "scft_engine_info",
    "scft_backward", "scft_spline_setup", "scft_spline_set_x", "scft_spline_set_xref", "scft_spline_derivs",
    "scft_spline_redist", "scft_spline_weights", "scft_spline_derivs_noalloc", "scft_autotune",
]
_STRING_RESULT = ("scft_last_error", "scft_fft_path", "scft_engine_info")


class ScftError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"scft status {code}: {msg}")
        self.code = code


(SCFT_SCHED_DEFAULT, SCFT_SCHED_TWO_STAGE, SCFT_SCHED_STOCKHAM, SCFT_SCHED_PENCIL, SCFT_SCHED_TWO_PASS,
 SCFT_SCHED_PENCIL_SPLIT) = range(6)
_FAMILY = {"default": SCFT_SCHED_DEFAULT, "two-stage": SCFT_SCHED_TWO_STAGE, "stockham": SCFT_SCHED_STOCKHAM,
           "pencil": SCFT_SCHED_PENCIL, "two-pass": SCFT_SCHED_TWO_PASS, "pencil-split": SCFT_SCHED_PENCIL_SPLIT}


class Schedule(C.Structure):
    """scft_schedule: kernel family, two-stage column tile and plane workgroup size,
    replica-group streams (0 in a field: the default)."""
    _fields_ = [("family", C.c_int), ("coltile", C.c_int), ("plane_nt", C.c_int), ("streams", C.c_int)]

    def as_dict(self):
        fam = {v: k for k, v in _FAMILY.items()}[self.family]
        return {"family": fam, "coltile": self.coltile, "plane_nt": self.plane_nt, "streams": self.streams}


def make_schedule(sched):
    """Schedule from None, a Schedule, or a dict with any of family ("two-stage", "two-pass",
    "stockham", "pencil", "pencil-split"), coltile, plane_nt, streams."""
    if sched is None:
        return Schedule()
    if isinstance(sched, Schedule):
        return sched
    fam = sched.get("family", "default")
    return Schedule(family=_FAMILY[fam] if isinstance(fam, str) else int(fam), coltile=int(sched.get("coltile", 0)),
                    plane_nt=int(sched.get("plane_nt", 0)), streams=int(sched.get("streams", 0)))


class Params(C.Structure):
    _fields_ = [
        ("m", C.c_int * 3), ("D", C.c_double * 3), ("N", C.c_int), ("f", C.c_double),
        ("alpha", C.c_double), ("ens", C.c_int), ("phidb", C.c_double), ("chi", C.c_double),
        ("nrep", C.c_int), ("precision", C.c_int), ("fix", C.c_int * 2), ("dostring", C.c_int),
        ("justFE", C.c_int), ("springsteps", C.c_int), ("device", C.c_int), ("fe_fp64", C.c_int),
        ("phihA", C.c_double), ("anderson", C.c_int), ("anderson_lambda", C.c_double),
        ("anderson_start", C.c_double), ("autotune", C.c_int), ("ckpt", C.c_int), ("sched", Schedule),
    ]


class IterReport(C.Structure):
    _fields_ = [
        ("errs", C.POINTER(C.c_double)), ("errs3", C.POINTER(C.c_double)), ("x", C.POINTER(C.c_double)),
        ("xref", C.POINTER(C.c_double)), ("dotWs", C.POINTER(C.c_double)), ("dotWs2", C.POINTER(C.c_double)),
        ("normDER", C.POINTER(C.c_double)), ("normDEV", C.POINTER(C.c_double)), ("steps", C.POINTER(C.c_int)),
        ("fitup", C.c_double), ("stepav", C.c_double),
    ]


_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built; run `python -m scft_string_amd.build`")
        L = C.CDLL(LIB_PATH)
        vp, dp, ip = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int)
        L.scft_ctx_create.argtypes = [C.POINTER(Params), C.POINTER(vp)]
        L.scft_ctx_destroy.argtypes = [vp]
        L.scft_last_error.argtypes = [vp]
        L.scft_last_error.restype = C.c_char_p
        L.scft_sync.argtypes = [vp]
        L.scft_props_batch.argtypes = [vp, dp, dp, dp, dp, C.c_int]
        L.scft_props_batch_device.argtypes = [vp, C.c_void_p, C.c_void_p, dp, dp, C.c_int]
        L.scft_set_fields.argtypes = [vp, dp]
        L.scft_get_fields.argtypes = [vp, dp]
        L.scft_get_densities.argtypes = [vp, dp]
        L.scft_get_tangent.argtypes = [vp, dp]
        L.scft_solve_w_plus.argtypes = [vp, C.c_int, C.c_double, ip, ip]
        L.scft_string_iterate.argtypes = [vp, C.c_int, C.c_double, dp, C.POINTER(IterReport)]
        L.scft_free_energy.argtypes = [vp, dp, dp, dp, dp, dp, dp]
        L.scft_debug_qq_invariant.argtypes = [vp, C.c_int, dp]
        L.scft_debug_time_props.argtypes = [vp, C.c_int, C.c_int, dp]
        L.scft_fft_path.argtypes = [vp]
        L.scft_fft_path.restype = C.c_char_p
        L.scft_engine_info.argtypes = [vp]
        L.scft_engine_info.restype = C.c_char_p
        L.scft_autotune.argtypes = [C.POINTER(Params), dp, C.c_int, C.POINTER(Schedule), C.c_char_p, C.c_size_t]
        L.scft_dlnq_dbox.argtypes = [vp, dp, C.c_int, dp, dp]
        L.scft_box_stress.argtypes = [vp, dp, dp]
        L.scft_backward.argtypes = [vp, dp, dp, dp]
        L.scft_set_box.argtypes = [vp, dp]
        L.scft_relax_box.argtypes = [vp, C.c_int, C.c_double, C.c_double, C.c_int, C.c_int, dp, dp, ip]
        L.scft_iterate_concurrent.argtypes = [C.POINTER(vp), C.c_int, ip, C.c_double, C.POINTER(dp),
                                              C.POINTER(IterReport)]
        L.scft_spline_setup.argtypes = [vp, C.c_int]
        L.scft_spline_set_x.argtypes = [vp, dp]
        L.scft_spline_set_xref.argtypes = [vp, dp]
        L.scft_spline_derivs.argtypes = [vp, C.POINTER(dp), C.c_int, C.c_int, ip, C.POINTER(dp)]
        L.scft_spline_redist.argtypes = [vp, C.POINTER(dp), C.c_int, C.c_int, ip, C.c_double]
        L.scft_spline_weights.argtypes = [C.c_int, dp, dp, ip, C.c_int, C.c_int, dp]
        L.scft_spline_derivs_noalloc.argtypes = [C.c_int, C.c_longlong, dp, ip, C.c_void_p, C.c_longlong,
                                                 C.c_void_p, C.c_longlong, C.c_void_p]
        for name in EXPORTED:
            if name not in _STRING_RESULT:
                getattr(L, name).restype = C.c_int
        _LIB = L
    return _LIB


def _dp(a):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _rows(arr):
    """double** over the rows of a C-contiguous float64 2-D array"""
    dp = C.POINTER(C.c_double)
    return (dp * arr.shape[0])(*[arr[i].ctypes.data_as(dp) for i in range(arr.shape[0])])


def spline_weights(x, xe, erows, deriv):
    """scft_spline_weights: (len(erows), len(x)) weights of the natural spline through x at
    xe[erows] (derivative or value). Host only."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    xe = np.ascontiguousarray(xe, dtype=np.float64)
    er = np.ascontiguousarray(erows, dtype=np.int32)
    G = np.zeros((len(er), len(x)))
    rc = lib().scft_spline_weights(len(x), _dp(x), _dp(xe), er.ctypes.data_as(C.POINTER(C.c_int)), len(er),
                                   int(bool(deriv)), _dp(G))
    if rc != SCFT_OK:
        raise ScftError(rc, "scft_spline_weights: bad arguments")
    return G


def autotune(m, D, N, f, alpha, ens, phidb, chi, nrep=1, precision="fp64", fix=(0, 0), dostring=1, W=None, reps=2,
             ckpt=0, device=0):
    """scft_autotune: times every schedule the grid can run, each in a context of its own, on
    W ((nrep, 2, M); None: a lamellar trial field). Returns (schedule dict, report text)."""
    prec = {"fp64": SCFT_FP64, "fp32": SCFT_FP32}[precision]
    p = Params(m=(C.c_int * 3)(*[int(v) for v in m]), D=(C.c_double * 3)(*[float(v) for v in D]), N=int(N),
               f=float(f), alpha=float(alpha), ens=int(ens), phidb=float(phidb), chi=float(chi), nrep=int(nrep),
               precision=prec, fix=(C.c_int * 2)(*fix), dostring=int(dostring), device=int(device), fe_fp64=1,
               ckpt=int(ckpt))
    Wp = None
    if W is not None:
        W = np.ascontiguousarray(W, dtype=np.float64).reshape(int(nrep), -1)
        Wp = _dp(W)
    best = Schedule()
    buf = C.create_string_buffer(8192)
    rc = lib().scft_autotune(C.byref(p), Wp, int(reps), C.byref(best), buf, len(buf))
    if rc != SCFT_OK:
        raise ScftError(rc, "scft_autotune failed")
    return best.as_dict(), buf.value.decode()


class Context:
    """One scft_ctx: grid, chain parameters, nrep replicas' device state. schedule: None (the
    defaults), a dict as returned by autotune(), or a Schedule."""

    def __init__(self, m, D, N, f, alpha, ens, phidb, chi, nrep=1, precision="fp64", fix=(0, 0),
                 dostring=1, justFE=0, springsteps=0, device=0, fe_fp64=1, phihA=0.0, anderson=0,
                 anderson_lambda=0.0, anderson_start=0.0, autotune=0, ckpt=0, schedule=None):
        self.m = tuple(int(v) for v in m)
        self.D = tuple(float(v) for v in D)
        self.M = int(np.prod(self.m))
        self.nrep = int(nrep)
        self.N = int(N)
        self.device = int(device)
        prec = {"fp64": SCFT_FP64, "fp32": SCFT_FP32}[precision]
        p = Params(m=(C.c_int * 3)(*self.m), D=(C.c_double * 3)(*[float(v) for v in D]), N=int(N), f=float(f),
                   alpha=float(alpha), ens=int(ens), phidb=float(phidb), chi=float(chi), nrep=self.nrep,
                   precision=prec, fix=(C.c_int * 2)(*fix), dostring=int(dostring), justFE=int(justFE),
                   springsteps=int(springsteps), device=int(device), fe_fp64=int(fe_fp64),
                   phihA=float(phihA), anderson=int(anderson), anderson_lambda=float(anderson_lambda),
                   anderson_start=float(anderson_start), autotune=int(autotune), ckpt=int(ckpt),
                   sched=make_schedule(schedule))
        self.params = p
        h = C.c_void_p()
        rc = lib().scft_ctx_create(C.byref(p), C.byref(h))
        if rc != SCFT_OK:
            raise ScftError(rc, f"scft_ctx_create failed for grid {self.m} ({precision})")
        self._h = h

    def _chk(self, rc):
        if rc != SCFT_OK:
            raise ScftError(rc, lib().scft_last_error(self._h).decode())

    def close(self):
        if getattr(self, "_h", None):
            lib().scft_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def fft_path(self):
        """The propagator's contour-step kernels for this grid: "two-pass", "two-stage" (with
        "-c32" / "-c64" appended at the family's other column tile width), "stockham",
        "pencil-split" or "pencil"."""
        return lib().scft_fft_path(self._h).decode()

    @property
    def engine_info(self):
        """The propagator configuration, with the autotuner's timings when it ran."""
        return lib().scft_engine_info(self._h).decode()

    # ---------------------------------------------------------------- props (density.h:344)
    def props_batch(self, W):
        """W: (n, 2, M) or (n, 2M). Returns phi (n, 4, M), lnQ (n,), phis (n, 5)."""
        W = np.ascontiguousarray(W, dtype=np.float64).reshape(-1, 2 * self.M)
        n = W.shape[0]
        phi = np.empty((n, 4 * self.M))
        lnQ = np.empty(n)
        phis = np.empty((n, 5))
        self._chk(lib().scft_props_batch(self._h, _dp(W), _dp(phi), _dp(lnQ), _dp(phis), n))
        return phi.reshape(n, 4, self.M), lnQ, phis

    def props_batch_device(self, W_dev, phi_dev, n):
        """Device-pointer variant: W_dev (n rows of 2M) and phi_dev (n rows of 4M) are torch
        float64 tensors on the context's GPU, C-contiguous. The library reads W_dev on its own
        stream, so the producer's stream (torch's current stream) is synchronised first; phi_dev
        is complete when this returns."""
        import torch

        n = int(n)
        if not 0 < n <= self.nrep:
            raise ValueError(f"props_batch_device: n = {n} outside 1..{self.nrep}")
        for name, t, need in (("W_dev", W_dev, 2 * self.M * n), ("phi_dev", phi_dev, 4 * self.M * n)):
            if not isinstance(t, torch.Tensor) or t.device.type != "cuda" or t.device.index != self.device:
                raise ValueError(f"props_batch_device: {name} must be a tensor on cuda:{self.device}")
            if t.dtype != torch.float64 or not t.is_contiguous():
                raise ValueError(f"props_batch_device: {name} must be contiguous float64")
            if t.numel() < need:
                raise ValueError(f"props_batch_device: {name} holds {t.numel()} values, needs {need}")
        torch.cuda.current_stream(W_dev.device).synchronize()
        lnQ = np.empty(n)
        phis = np.empty((n, 5))
        self._chk(lib().scft_props_batch_device(self._h, C.c_void_p(W_dev.data_ptr()),
                                                C.c_void_p(phi_dev.data_ptr()), _dp(lnQ), _dp(phis), n))
        return lnQ, phis

    # ------------------------------------------------ chunked host-row spline (density.h:462-528)
    def spline_setup(self, chunk):
        self._chk(lib().scft_spline_setup(self._h, int(chunk)))

    def spline_set_x(self, x):
        self._chk(lib().scft_spline_set_x(self._h, _dp(np.ascontiguousarray(x, dtype=np.float64))))

    def spline_set_xref(self, xref):
        self._chk(lib().scft_spline_set_xref(self._h, _dp(np.ascontiguousarray(xref, dtype=np.float64))))

    def spline_derivs(self, ww, r0, n, doiis, out):
        """ww, out: (nrep, >= r0 + n) float64 host rows; out rows with doiis[ii] == 1 get the
        spline derivative at the nodes over points r0 .. r0+n-1."""
        assert ww.dtype == np.float64 and ww.flags["C_CONTIGUOUS"] and out.flags["C_CONTIGUOUS"]
        d = np.ascontiguousarray(doiis, dtype=np.int32)
        self._chk(lib().scft_spline_derivs(self._h, _rows(ww), int(r0), int(n), d.ctypes.data_as(C.POINTER(C.c_int)),
                                           _rows(out)))

    def spline_redist(self, ww, r0, n, doiis, lam):
        assert ww.dtype == np.float64 and ww.flags["C_CONTIGUOUS"]
        d = np.ascontiguousarray(doiis, dtype=np.int32)
        self._chk(lib().scft_spline_redist(self._h, _rows(ww), int(r0), int(n), d.ctypes.data_as(C.POINTER(C.c_int)),
                                           float(lam)))

    def qq_invariant(self, slot=0):
        out = np.empty(self.N + 1)
        self._chk(lib().scft_debug_qq_invariant(self._h, int(slot), _dp(out)))
        return out

    def time_props(self, nrep, reps=3):
        """mean ms per contour step of props over the first nrep field-state rows"""
        out = C.c_double()
        self._chk(lib().scft_debug_time_props(self._h, int(nrep), int(reps), C.byref(out)))
        return out.value

    # ---------------------------------------------------------------- field state
    def set_fields(self, W):
        W = np.ascontiguousarray(W, dtype=np.float64).reshape(self.nrep, 2 * self.M)
        self._chk(lib().scft_set_fields(self._h, _dp(W)))

    def get_fields(self):
        W = np.empty((self.nrep, 2 * self.M))
        self._chk(lib().scft_get_fields(self._h, _dp(W)))
        return W

    def get_densities(self):
        phi = np.empty((self.nrep, 4 * self.M))
        self._chk(lib().scft_get_densities(self._h, _dp(phi)))
        return phi

    def get_tangent(self):
        t = np.empty((self.nrep, self.M))
        self._chk(lib().scft_get_tangent(self._h, _dp(t)))
        return t

    # ---------------------------------------------------------------- solvers (scft.cu)
    def solve_w_plus(self, maxIter, errTol, which=None):
        k = np.zeros(self.nrep, dtype=np.int32)
        wp = None
        if which is not None:
            which = np.ascontiguousarray(which, dtype=np.int32)
            wp = which.ctypes.data_as(C.POINTER(C.c_int))
        self._chk(lib().scft_solve_w_plus(self._h, int(maxIter), float(errTol), wp,
                                          k.ctypes.data_as(C.POINTER(C.c_int))))
        return k

    def _report(self):
        R = self.nrep
        arrs = {n: np.zeros(R) for n in ("errs", "errs3", "x", "xref", "dotWs", "dotWs2", "normDER", "normDEV")}
        steps = np.zeros(R, dtype=np.int32)
        alf = np.zeros(R)
        rep = IterReport(**{n: _dp(a) for n, a in arrs.items()}, steps=steps.ctypes.data_as(C.POINTER(C.c_int)))
        return arrs, steps, alf, rep

    @staticmethod
    def _result(arrs, steps, alf, rep):
        out = dict(arrs)
        out["steps"] = steps
        out["alf"] = alf
        out["fitup"] = rep.fitup
        out["stepav"] = rep.stepav
        return out

    def string_iterate(self, k, errTol=1e-3):
        arrs, steps, alf, rep = self._report()
        self._chk(lib().scft_string_iterate(self._h, int(k), float(errTol), _dp(alf), C.byref(rep)))
        return self._result(arrs, steps, alf, rep)

    def free_energy(self):
        R = self.nrep
        FE = C.c_double()
        FEs, FEs0, FEs1, ph, alf = (np.zeros(R) for _ in range(5))
        self._chk(lib().scft_free_energy(self._h, C.byref(FE), _dp(FEs), _dp(FEs0), _dp(FEs1), _dp(ph), _dp(alf)))
        return {"FE": FE.value, "FEs": FEs, "FEs0": FEs0, "FEs1": FEs1, "phctot": ph, "alf": alf}

    # ---------------------------------------------------------------- box derivatives
    def dlnq_dbox(self, W):
        """d lnQ / dD_i at fixed W for rows W (n, 2, M): returns lnQ (n,), dlnQ_dD (n, 3)."""
        W = np.ascontiguousarray(W, dtype=np.float64).reshape(-1, 2 * self.M)
        n = W.shape[0]
        lnQ = np.empty(n)
        d = np.empty((n, 3))
        self._chk(lib().scft_dlnq_dbox(self._h, _dp(W), n, _dp(lnQ), _dp(d)))
        return lnQ, d

    def box_stress(self):
        """dF/dD (nrep, 3) and dF/dchiN (nrep,) at the current field state."""
        dFdD = np.empty((self.nrep, 3))
        dFdchi = np.empty(self.nrep)
        self._chk(lib().scft_box_stress(self._h, _dp(dFdD), _dp(dFdchi)))
        return dFdD, dFdchi

    def backward(self, dL_dF):
        """scft_backward: dL/dD (nrep, 3) and dL/dchiN (nrep,) from dL/dF (nrep,) at the
        current field state."""
        g = np.ascontiguousarray(dL_dF, dtype=np.float64).reshape(self.nrep)
        dD = np.empty((self.nrep, 3))
        dchi = np.empty(self.nrep)
        self._chk(lib().scft_backward(self._h, _dp(g), _dp(dD), _dp(dchi)))
        return dD, dchi

    def set_box(self, D):
        D = np.ascontiguousarray(D, dtype=np.float64).reshape(3)
        self._chk(lib().scft_set_box(self._h, _dp(D)))
        self.D = tuple(float(v) for v in D)

    def relax_box(self, axes=(1, 0, 0), stress_tol=1e-6, err_tol=1e-9, max_field_iter=5000, max_boxes=30):
        """Stress-free box relaxation (single replica, dostring = 0). Returns (D, dF/dD, boxes)."""
        mask = sum(1 << i for i in range(3) if axes[i])
        D = np.ascontiguousarray(self.D, dtype=np.float64)
        g = np.zeros(3)
        nbox = C.c_int()
        self._chk(lib().scft_relax_box(self._h, mask, float(stress_tol), float(err_tol), int(max_field_iter),
                                       int(max_boxes), _dp(D), _dp(g), C.byref(nbox)))
        self.D = tuple(float(v) for v in D)
        return self.D, g, nbox.value


def iterate_concurrent(ctxs, ks, errTol=1e-3):
    """One string_iterate on each context at once (scft_iterate_concurrent). Returns the
    per-context reports, as Context.string_iterate would."""
    n = len(ctxs)
    parts = [c._report() for c in ctxs]
    hs = (C.c_void_p * n)(*[c._h for c in ctxs])
    kk = (C.c_int * n)(*[int(k) for k in ks])
    alfs = (C.POINTER(C.c_double) * n)(*[_dp(p[2]) for p in parts])
    reps = (IterReport * n)(*[p[3] for p in parts])
    rc = lib().scft_iterate_concurrent(hs, n, kk, float(errTol), alfs, reps)
    if rc != SCFT_OK:
        for c in ctxs:
            msg = lib().scft_last_error(c._h).decode()
            if msg:
                raise ScftError(rc, msg)
        raise ScftError(rc, "scft_iterate_concurrent failed")
    return [Context._result(p[0], p[1], p[2], reps[i]) for i, p in enumerate(parts)]
