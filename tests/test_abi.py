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
