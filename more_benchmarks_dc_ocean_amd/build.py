"""Build the HIP library in-tree for gfx950.

``build()`` compiles ``csrc/redi_lib.hip`` (which includes every kernel
header) into ``more_benchmarks_dc_ocean_amd/libredi_hip.so`` with hipcc.
FP contraction is off so the FP64 kernels round like the reference build;
packed-FP32 instructions are disabled.  The library links ROCm's roctx
library for the named ranges around its public calls.
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
    cmd = [HIPCC, *hipcc_flags(), src, "-lrocprofiler-sdk-roctx", "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        sys.stderr.write(p.stdout + p.stderr)
        raise RuntimeError("hipcc failed building libredi_hip.so")
    os.replace(tmp, LIB)
    return LIB


RK2_LIB = os.path.join(PKG, "librk2_redi_hip.so")


def build_rk2(verbose: bool = False) -> str:
    """The RK2 benchmark seam (csrc/rk2_redi_hip.cpp, host code over
    libredi_hip.so) as a self-contained -Bsymbolic library, as the
    reference's RK2 driver links each kernel."""
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    cmd = [os.environ.get("CXX", "g++"), "-O2", "-std=c++17", "-fPIC", "-shared",
           "-D__HIP_PLATFORM_AMD__", "-I" + os.path.join(rocm, "include"),
           "-I" + os.path.join(ROOT, "include"), os.path.join(CSRC, "rk2_redi_hip.cpp"),
           "-L" + PKG, "-lredi_hip", "-L" + os.path.join(rocm, "lib"), "-lamdhip64",
           "-Wl,-Bsymbolic", "-Wl,-rpath,$ORIGIN", "-Wl,-rpath," + os.path.join(rocm, "lib"),
           "-o", RK2_LIB + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        sys.stderr.write(p.stdout + p.stderr)
        raise RuntimeError("building librk2_redi_hip.so failed")
    os.replace(RK2_LIB + ".tmp", RK2_LIB)
    return RK2_LIB


FORTRAN = os.path.join(PKG, "fortran")
FBUILD = os.path.join(FORTRAN, "_build")
DROPIN_SRC = os.path.join(ROOT, "tests", "fortran", "redi_dropin_driver.F90")
DROPIN_BIN = os.path.join(ROOT, "tests", "fortran", "_build", "redi_dropin_hip")
FC = os.environ.get("FC", "/opt/rocm/llvm/bin/amdflang")


def build_fortran(verbose: bool = False) -> str:
    """The Fortran drop-in modules (fortran/*.F90) and the reference-style
    test driver linked against them and libredi_hip.so; returns the driver
    path, or "" where no Fortran compiler is installed."""
    if not os.path.exists(FC):
        return ""
    os.makedirs(FBUILD, exist_ok=True)
    os.makedirs(os.path.dirname(DROPIN_BIN), exist_ok=True)
    flags = ["-O2", "-module-dir", FBUILD]
    objs = []
    for src in ("redi_hip_c.F90", "ocean_redi_hip.F90"):
        obj = os.path.join(FBUILD, src.replace(".F90", ".o"))
        cmd = [FC, *flags, "-c", os.path.join(FORTRAN, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True, capture_output=not verbose)
        objs.append(obj)
    drv = os.path.join(FBUILD, "redi_dropin_driver.o")
    subprocess.run([FC, *flags, "-c", DROPIN_SRC, "-o", drv], check=True, capture_output=not verbose)
    rocm_lib = os.environ.get("ROCM_LIB", "/opt/rocm/lib")
    cmd = [FC, "-o", DROPIN_BIN, *objs, drv, "-L" + PKG, "-lredi_hip", "-L" + rocm_lib,
           "-lamdhip64", "-Wl,-rpath,$ORIGIN/../../../more_benchmarks_dc_ocean_amd",
           "-Wl,-rpath," + rocm_lib]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True, capture_output=not verbose)
    return DROPIN_BIN


if __name__ == "__main__":
    print(build(verbose=True))
    print(build_rk2(verbose=True))
    print(build_fortran(verbose=True))
