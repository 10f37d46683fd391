"""Build libfipi.so in-tree with hipcc for gfx950.

Usage: python -m fipi_amd.build  (or fipi_amd.build.build() from Python)
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libfipi.so")
ARCH = os.environ.get("FIPI_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = [
    "--offload-arch=" + ARCH,
    "-O3",
    "-std=c++17",
    "-fPIC",
    # every multiply and add rounded separately, as in the reference build
    "-ffp-contract=off",
    # no packed FP32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build(verbose: bool = False) -> str:
    obj_init = os.path.join(CSRC, "init.o")
    _run(["g++", "-O2", "-fPIC", "-std=c++17", "-ffp-contract=off", "-c",
          os.path.join(CSRC, "init.cpp"), "-o", obj_init], verbose)
    obj_hip = os.path.join(CSRC, "fipi.o")
    _run(["hipcc", *HIP_FLAGS, "-c", "-o", obj_hip, os.path.join(CSRC, "fipi.hip")], verbose)
    tmp = LIB + ".tmp"
    _run(["hipcc", "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp, obj_hip, obj_init], verbose)
    os.replace(tmp, LIB)
    os.remove(obj_init)
    os.remove(obj_hip)
    return LIB


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(LIB)
