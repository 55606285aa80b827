"""Build libpsx_scft.so (HIP, gfx950) in-tree.

    python -m polyswift_plusplus_amd.build_ext

hipcc cross-compiles for gfx950 without a GPU.  The library lands next to
this file so it travels with the source tree.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libpsx_scft.so")
SOURCES = [os.path.join(CSRC, "psx_scft.hip")]
DEPS = SOURCES + [os.path.join(CSRC, f) for f in ("psx_kernels.hpp", "psx_fft_lds.hpp")] + [
    os.path.join(ROOT, "include", "psx_scft.h")]

HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    # packed FP32 instructions off for all device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found (ROCm required to build libpsx_scft.so)")


def build(verbose: bool = False) -> str:
    tmp = LIB + ".tmp"
    cmd = [hipcc(), *HIP_FLAGS, "-I", os.path.join(ROOT, "include"), "-o", tmp, *SOURCES]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    errs = "\n".join(l for l in (r.stderr or "").splitlines()
                     if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed ({r.returncode}):\n{errs}")
    if verbose and errs.strip():
        print(errs, file=sys.stderr)
    os.replace(tmp, LIB)
    build_demo(verbose)
    return LIB


DEMO = os.path.join(HERE, "psx_demo")


def build_demo(verbose: bool = False) -> str:
    """C++ host program over the C ABI (links libpsx_scft.so by rpath)."""
    cmd = ["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include"),
           os.path.join(CSRC, "psx_demo.cpp"), "-L", HERE, "-lpsx_scft",
           "-Wl,-rpath,$ORIGIN", "-o", DEMO]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"g++ failed ({r.returncode}):\n{r.stderr}")
    return DEMO


if __name__ == "__main__":
    print(build(verbose=True))
