"""The C-ABI library loads and exports every entry point declared in include/*.h
(no GPU calls)."""
import glob
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = []
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        txt = open(h).read()
        names += re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(scft_\w+)\s*\(", txt, flags=re.M)
    return sorted(set(names))


def test_header_declares_entry_points():
    names = declared_functions()
    assert "scft_ctx_create" in names and "scft_props_batch" in names and len(names) >= 10


def test_library_exports_every_declared_symbol():
    import ctypes

    from scft_string_amd import capi

    if not os.path.exists(capi.LIB_PATH):
        pytest.skip("libscft_hip.so not built")
    L = ctypes.CDLL(capi.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing
    assert set(capi.EXPORTED) == set(declared_functions())


def test_python_layer_fails_loudly_without_library(monkeypatch, tmp_path):
    from scft_string_amd import capi

    monkeypatch.setattr(capi, "LIB_PATH", str(tmp_path / "nope.so"))
    monkeypatch.setattr(capi, "_LIB", None)
    with pytest.raises(ImportError):
        capi.lib()


def test_autotune_is_refused_before_any_device_work():
    """autotune = 1 is switched off: scft_ctx_create refuses it while validating the
    parameters, before it selects a device (so this runs without a GPU)."""
    import ctypes as C

    from scft_string_amd import capi

    if not os.path.exists(capi.LIB_PATH):
        pytest.skip("libscft_hip.so not built")
    p = capi.Params(m=(C.c_int * 3)(16, 16, 16), D=(C.c_double * 3)(4.0, 4.0, 4.0), N=20, f=0.5, alpha=0.1, ens=1,
                    phidb=0.8, chi=15.0, nrep=1, precision=capi.SCFT_FP64, fix=(C.c_int * 2)(0, 0), dostring=0,
                    autotune=1)
    h = C.c_void_p()
    assert capi.lib().scft_ctx_create(C.byref(p), C.byref(h)) == capi.SCFT_ERR_UNSUPPORTED
    assert not h.value
