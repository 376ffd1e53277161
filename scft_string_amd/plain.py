"""ctypes binding of libscft_plain.so: the benchmark's comparator, the plain per-replica FP64
GPU path (hipFFT D2Z/Z2D, one launch per elementwise op, blocking copies per props). It is a
baseline to measure against, not the product path."""
import ctypes as C
import os

import numpy as np

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libscft_plain.so")
_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built; run `python -m scft_string_amd.build`")
        L = C.CDLL(LIB_PATH)
        dp = C.POINTER(C.c_double)
        L.plain_create.restype = C.c_void_p
        L.plain_create.argtypes = [C.POINTER(C.c_int), dp]
        L.plain_destroy.argtypes = [C.c_void_p]
        L.plain_props.restype = C.c_int
        L.plain_props.argtypes = [C.c_void_p, dp, dp, C.c_int, C.c_double, C.c_double, C.c_double, C.c_int]
        _LIB = L
    return _LIB


class PlainFP64:
    def __init__(self, m, D, N):
        self.M = int(np.prod(m))
        self._h = lib().plain_create((C.c_int * 3)(*m), (C.c_double * 3)(*D), int(N))
        if not self._h:
            raise RuntimeError("plain_create failed")

    def props(self, W, N, alpha, phic, f, ens):
        w = np.zeros(6 * self.M)
        w[:2 * self.M] = np.asarray(W, dtype=np.float64).reshape(-1)
        lnQ = C.c_double()
        rc = lib().plain_props(self._h, w.ctypes.data_as(C.POINTER(C.c_double)), C.byref(lnQ), int(N),
                               float(alpha), float(phic), float(f), int(ens))
        if rc != 0:
            raise RuntimeError(f"plain_props failed ({rc})")
        return w[2 * self.M:].reshape(4, self.M), lnQ.value

    def close(self):
        if self._h:
            lib().plain_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
