"""Build the native libraries in-tree (no JIT cache, so they travel with the
source tree):

* cable_amd/libcable_gpu.so  -- HIP kernels and C ABI (include/cable_gpu.h),
  compiled for gfx950 with hipcc;
* oracle/liboracle.so, oracle/_ref/*  -- the test oracle and, when the
  reference tree is present, the compiled reference (oracle/Makefile).
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "cable_amd", "csrc")
LIB = os.path.join(ROOT, "cable_amd", "libcable_gpu.so")

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    # no FMA contraction: the FP64 path reproduces the reference's rounding
    "-ffp-contract=off",
    # no packed FP32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def hipcc():
    for c in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def build_gpu(verbose=False):
    src = os.path.join(CSRC, "cable_gpu.hip")
    tmp = LIB + ".tmp"
    cmd = [hipcc()] + HIPCC_FLAGS + ["-o", tmp, src]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed for libcable_gpu.so")
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


def build_oracle(with_reference=True, verbose=False):
    targets = ["oracle"]
    if with_reference and os.path.isdir("/root/reference/src"):
        # the compiled reference, and the reference driven through the
        # Fortran drop-in for cbm (needs libcable_gpu.so built first)
        targets += ["r8", "r4", "gpu8", "gpu4"]
    r = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")] + targets,
                       capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout[-4000:] + r.stderr[-4000:])
        raise RuntimeError("oracle build failed")
    if verbose:
        print("built oracle targets:", " ".join(targets))


def build_all(verbose=True):
    build_gpu(verbose)
    build_oracle(verbose=verbose)


if __name__ == "__main__":
    build_all()
