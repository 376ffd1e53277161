"""Builds libscft_hip.so (the HIP product library) in-tree for gfx950.

    python -m scft_string_amd.build          # library only
    python -m scft_string_amd.build --driver # also the C++ host driver scft_drive
"""
import argparse
import concurrent.futures as cf
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INC = os.path.join(ROOT, "include")
OBJ = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libscft_hip.so")
DRIVER = os.path.join(PKG, "scft_drive")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("SCFT_OFFLOAD_ARCH", "gfx950")
# Packed-FP32 instructions are disabled for all device code.
COMMON = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-Xclang", "-target-feature",
          "-Xclang", "-packed-fp32-ops", "-I", INC, "-I", CSRC, "-Wall", "-Wno-unused-result"]

SOURCES = ["prop_engine.hip", "field_kernels.hip", "scft_api.hip"]


def _compile(src):
    out = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, *COMMON, "-c", os.path.join(CSRC, src), "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    err = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l and l.strip())
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{err}")
    return out, err


def build_library(verbose=True):
    os.makedirs(OBJ, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(4, len(SOURCES))) as ex:
        results = list(ex.map(_compile, SOURCES))
    objs = [o for o, _ in results]
    for _, warn in results:
        if warn and verbose:
            print(warn, file=sys.stderr)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


def build_driver(verbose=True):
    src = os.path.join(CSRC, "driver", "scft_drive.cpp")
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-I", INC, "-I", os.path.join(CSRC, "driver"), src, "-o", DRIVER,
           f"-L{PKG}", "-lscft_hip", f"-Wl,-rpath,$ORIGIN", "-Wl,-rpath,/opt/rocm/lib"]
    subprocess.run(cmd, check=True)
    if verbose:
        print("built", DRIVER)
    return DRIVER


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--driver", action="store_true")
    a = ap.parse_args()
    build_library()
    if a.driver:
        build_driver()


if __name__ == "__main__":
    main()
