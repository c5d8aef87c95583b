"""Compile the HIP library (gfx950) in-tree: fsilbm3d_amd/libfsil.so.

Flags:
  --offload-arch=gfx950            MI355X (CDNA4) code objects only
  -ffp-contract=off                no fused multiply-add: the FP64 form rounds
                                   every product and sum as the reference does
  -target-feature -packed-fp32-ops no packed FP32 instructions
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")
LIB = os.path.join(HERE, "libfsil.so")
SOURCES = ["fsil_capi.hip"]
HEADERS = ["fsil_common.h", "fsil_lbm_kernels.h", "fsil_lbm32_kernels.h", "fsil_ibm_kernels.h", "fsil_beam.h",
           "fsil_fem_kernels.h", "fsil_tree_kernels.h", "fsil_mlp_kernels.h"]


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def build_library(verbose: bool = False) -> str:
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
           "-ffp-contract=off", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
           "-Wno-unused-command-line-argument",
           "-I", CSRC, "-I", INCLUDE, "-o", tmp] + [os.path.join(CSRC, s) for s in SOURCES]
    if verbose:
        print(" ".join(cmd), flush=True)
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + p.stdout[-6000:] + p.stderr[-6000:])
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build_library(verbose="-v" in sys.argv))
