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
DROPIN = os.path.join(PKG, "scft_dropin_example")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("SCFT_OFFLOAD_ARCH", "gfx950")
# Packed-FP32 instructions are disabled for all device code.
COMMON = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-Xclang", "-target-feature",
          "-Xclang", "-packed-fp32-ops", "-I", INC, "-I", CSRC, "-Wall", "-Wno-unused-result"]

SOURCES = ["prop_engine.hip", "field_kernels.hip", "scft_api.hip", "scft_noalloc.hip"]


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


PLAIN_LIB = os.path.join(PKG, "libscft_plain.so")


def build_plain(verbose=True):
    """The benchmark's comparator: the plain per-replica FP64 path (hipFFT, unfused)."""
    src = os.path.join(CSRC, "baseline", "plain_fp64.hip")
    tmp = PLAIN_LIB + ".tmp"
    cmd = [HIPCC, *COMMON, "-shared", src, "-o", tmp, "-L/opt/rocm/lib", "-lhipfft", "-Wl,-rpath,/opt/rocm/lib"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed for plain_fp64.hip:\n" + r.stderr)
    os.replace(tmp, PLAIN_LIB)
    if verbose:
        print("built", PLAIN_LIB)
    return PLAIN_LIB


def build_driver(verbose=True):
    """The C++ host driver (main() of scft.cu) and the drop-in example (scft.cu-shaped caller
    of the density mirror), both plain g++ against the C-ABI."""
    for name, out in (("scft_drive.cpp", DRIVER), ("dropin_example.cpp", DROPIN)):
        src = os.path.join(CSRC, "driver", name)
        cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-I", INC, "-I", os.path.join(CSRC, "driver"), src, "-o", out,
               f"-L{PKG}", "-lscft_hip", f"-Wl,-rpath,$ORIGIN", "-Wl,-rpath,/opt/rocm/lib"]
        subprocess.run(cmd, check=True)
        if verbose:
            print("built", out)
    return DRIVER


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--driver", action="store_true")
    a = ap.parse_args()
    build_library()
    build_plain()
    if a.driver:
        build_driver()


if __name__ == "__main__":
    main()
