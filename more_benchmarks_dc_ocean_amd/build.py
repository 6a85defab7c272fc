"""Build the HIP library in-tree for gfx950.

``build()`` compiles ``csrc/redi_lib.hip`` (which includes every kernel
header) into ``more_benchmarks_dc_ocean_amd/libredi_hip.so`` with hipcc.
FP contraction is off so the FP64 kernels round like the reference build;
packed-FP32 instructions are disabled.
"""
from __future__ import annotations

import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(PKG, "csrc")
ROOT = os.path.dirname(PKG)
LIB = os.path.join(PKG, "libredi_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("REDI_OFFLOAD_ARCH", "gfx950")


def hipcc_flags() -> list:
    return ["--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC", "-shared",
            "-ffp-contract=off", "-fno-fast-math",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-I" + os.path.join(ROOT, "include"), "-Wall", "-Wno-unused-function"]


def build(verbose: bool = False) -> str:
    src = os.path.join(CSRC, "redi_lib.hip")
    tmp = LIB + ".tmp"
    cmd = [HIPCC, *hipcc_flags(), src, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        sys.stderr.write(p.stdout + p.stderr)
        raise RuntimeError("hipcc failed building libredi_hip.so")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
