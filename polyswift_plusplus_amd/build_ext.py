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
DEPS = SOURCES + [os.path.join(CSRC, f) for f in ("psx_kernels.hpp", "psx_fft_lds.hpp", "psx_fft_reg.hpp",
                                                  "psx_fft_cline.hpp", "psx_fft_plane.hpp", "psx_filter.hpp",
                                                  "psx_rng.hpp")] + [
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
    build_h5(verbose)
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


H5LIB = os.path.join(HERE, "libpsx_h5.so")


def hdf5_root():
    """Install prefix of an HDF5 1.10+ C library (include/hdf5.h, lib/libhdf5.so)."""
    for cand in (os.environ.get("PSX_HDF5_ROOT"), "/opt/conda", "/usr"):
        if cand and os.path.exists(os.path.join(cand, "include", "hdf5.h")) and \
                os.path.exists(os.path.join(cand, "lib", "libhdf5.so")):
            return cand
    return None


def build_h5(verbose: bool = False):
    """libpsx_h5.so: PhysField dump/restore files (host C++ over HDF5).
    Skipped, with a message, when no HDF5 library is installed."""
    root = hdf5_root()
    if root is None:
        print("libpsx_h5.so not built: no HDF5 library found (set PSX_HDF5_ROOT)", file=sys.stderr)
        return None
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-I", os.path.join(root, "include"),
           os.path.join(CSRC, "psx_h5.cpp"), "-L", os.path.join(root, "lib"), "-lhdf5",
           "-Wl,-rpath," + os.path.join(root, "lib"), "-o", H5LIB]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"g++ failed ({r.returncode}):\n{r.stderr}")
    return H5LIB


if __name__ == "__main__":
    print(build(verbose=True))
