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


def test_bad_schedule_is_refused_before_any_device_work():
    """scft_ctx_create checks the schedule's ranges while it validates the parameters, before
    it selects a device (so this runs without a GPU)."""
    import ctypes as C

    from scft_string_amd import capi

    if not os.path.exists(capi.LIB_PATH):
        pytest.skip("libscft_hip.so not built")
    for sched in ({"family": 7}, {"streams": 5}, {"coltile": -1}, {"plane_nt": -64}):
        p = capi.Params(m=(C.c_int * 3)(16, 16, 16), D=(C.c_double * 3)(4.0, 4.0, 4.0), N=20, f=0.5, alpha=0.1,
                        ens=1, phidb=0.8, chi=15.0, nrep=1, precision=capi.SCFT_FP64, fix=(C.c_int * 2)(0, 0),
                        dostring=0, sched=capi.make_schedule(sched))
        h = C.c_void_p()
        assert capi.lib().scft_ctx_create(C.byref(p), C.byref(h)) == capi.SCFT_ERR_INVALID, sched
        assert not h.value


def test_ctypes_params_layout_matches_header(tmp_path):
    """capi.Params and capi.Schedule have the C header's sizes and field offsets (gcc on a
    probe of include/scft_api.h)."""
    import ctypes as C
    import shutil
    import subprocess

    from scft_string_amd import capi

    if not shutil.which("gcc"):
        pytest.skip("no gcc")
    fields = [f for f, _ in capi.Params._fields_]
    src = tmp_path / "probe.c"
    src.write_text("#include <stdio.h>\n#include <stddef.h>\n#include \"scft_api.h\"\nint main(void){\n"
                   'printf("%zu %zu\\n", sizeof(scft_params), sizeof(scft_schedule));\n'
                   + "".join(f'printf("%zu\\n", offsetof(scft_params, {f}));\n' for f in fields)
                   + "return 0;}\n")
    exe = tmp_path / "probe"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()
    assert int(out[0]) == C.sizeof(capi.Params) and int(out[1]) == C.sizeof(capi.Schedule)
    assert [int(v) for v in out[2:]] == [getattr(capi.Params, f).offset for f in fields]
