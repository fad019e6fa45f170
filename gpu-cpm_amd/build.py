"""Build libgpucpm.so (HIP kernels for gfx950 + C++ host engine + C ABI) in-tree.

Usage: python gpu-cpm_amd/build.py
The library lands next to the Python package (gpu-cpm_amd/gpucpm/), so it
travels with the source tree and is the file the package loads.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "gpucpm", "libgpucpm.so")
ARCH = os.environ.get("GPUCPM_ARCH", "gfx950")

SOURCES = ["kernels.hip", "engine.cpp"]

HIPCC_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "-shared",
    f"--offload-arch={ARCH}",
    # packed FP32 instructions stay disabled in all HIP code built here
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # no contraction of floating-point multiply-adds: the preferred-direction
    # update is the same double arithmetic as the oracle's
    "-ffp-contract=off",
    "-Wall", "-Wno-unused-command-line-argument",
]


def build(verbose: bool = False) -> str:
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = OUT + ".tmp"
    cmd = [hipcc, *HIPCC_FLAGS, "-o", tmp, *srcs]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose="-q" not in sys.argv))
