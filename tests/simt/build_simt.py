"""Host build of the library's own sources for CPU tests (g++, ASan + UBSan).

The library (more_benchmarks_dc_ocean_amd/csrc) is HIP for gfx950.  This
script copies its sources into tests/simt/_build/<variant>/src with three
mechanical rewrites -- a launch ``k<<<g, b, lds, st>>>(args)`` becomes
``simt::launch(k, g, b, lds, st)(args)``, a dynamic LDS array
``extern __shared__ T x[];`` becomes a pointer to the launch's LDS buffer,
and every static LDS declaration is followed by ``simt::lds_region`` (its
contents undefined at the start of each workgroup) -- and compiles them against tests/simt/hip/hip_runtime.h (the fiber execution
model of tests/simt/simt.hpp) together with tests/simt/driver.cpp.

Variants:
  base   the sources as they are;
  nt1    a lone tracer runs the one-tracer instantiation of the Phase-B
         kernels (the product runs it through the two-tracer kernels).

    python tests/simt/build_simt.py [--variant base|nt1] [--opt 0|1|2] [--sanitizer asan|msan]

prints the path of the driver binary.  Builds are cached by a hash of every
input and flag.
"""
from __future__ import annotations

import argparse
import hashlib
import os
import re
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
CSRC = os.path.join(ROOT, "more_benchmarks_dc_ocean_amd", "csrc")
INCLUDE = os.path.join(ROOT, "include")
BUILD = os.path.join(HERE, "_build")
LIB_SOURCES = ("redi_lib.hip", "redi_kernels.hpp", "redi_batch_host.hpp",
               "redi_batch_kernels.hpp", "redi_column.hpp", "redi_eos.hpp", "redi_eos_roquet.inc")
# asan: g++, AddressSanitizer + UndefinedBehaviorSanitizer.  msan: clang++
# (ROCm's host compiler), MemorySanitizer: a branch, address or output that
# depends on an uninitialized value -- which a GPU register would hold as
# garbage -- is reported with the allocation it came from.
SANITIZERS = {
    "asan": ("g++", ["-fsanitize=address,undefined", "-fno-sanitize-recover=all"]),
    "msan": (os.environ.get("SIMT_CLANG", "/opt/rocm/lib/llvm/bin/clang++"),
             ["-fsanitize=memory", "-fsanitize-memory-track-origins=2"]),
}

# kernel-name [template args] <<< config >>> (
_LAUNCH = re.compile(r"([A-Za-z_]\w*(?:<[^<>;(){}]*>)?)\s*<<<\s*(.*?)\s*>>>\s*\(", re.S)
_DYN_LDS = re.compile(r"extern\s+__shared__\s+(\w+)\s+(\w+)\s*\[\s*\]\s*;")
# a static LDS declaration: __shared__ type name[extent]... ;
_LDS = re.compile(r"(?<!extern )__shared__\s+([^;=]+?)\s+(\w+)\s*((?:\[[^;\]]*\])*)\s*;")

# nt1: the lone-tracer group of the single-domain Phase B runs NT = 1
_NT1_EDITS = (
    ("const int kn = gn == 1 ? 2 : gn;", "const int kn = gn;"),
    ("      case 2: launch_group<F, 2>(",
     "      case 1: launch_group<F, 1>(p, tiles, h, ts, recon, ry, fx, fy, nbx, nby, ua, st, tm); "
     "break;\n      case 2: launch_group<F, 2>("),
)


def rewrite(text: str) -> tuple[str, int]:
    n = len(_LAUNCH.findall(text))
    text = _LAUNCH.sub(lambda m: f"::simt::launch({m.group(1)}, {m.group(2)})(", text)
    text = _DYN_LDS.sub(lambda m: f"{m.group(1)}* {m.group(2)} = ({m.group(1)}*)::simt::dynamic_lds();",
                        text)
    text = _LDS.sub(lambda m: f"__shared__ {m.group(1)} {m.group(2)}{m.group(3)}; "
                              f"::simt::lds_region(&{m.group(2)}, sizeof({m.group(2)}));", text)
    return text, n


def prepare(variant: str, dst: str) -> int:
    os.makedirs(dst, exist_ok=True)
    launches = 0
    for name in LIB_SOURCES:
        with open(os.path.join(CSRC, name)) as f:
            text = f.read()
        text, n = rewrite(text)
        launches += n
        if name == "redi_lib.hip" and variant == "nt1":
            for old, new in _NT1_EDITS:
                if text.count(old) != 1:
                    raise RuntimeError(f"nt1 variant: {old!r} not found once in {name}")
                text = text.replace(old, new)
        out = "redi_lib.cpp" if name.endswith(".hip") else name
        with open(os.path.join(dst, out), "w") as f:
            f.write(text)
    if "<<<" in open(os.path.join(dst, "redi_lib.cpp")).read():
        raise RuntimeError("a kernel launch was not rewritten")
    return launches


def flags(opt: int, sanitizer: str = "asan") -> list:
    return ["-std=c++17", f"-O{opt}", "-g", "-fno-omit-frame-pointer", "-ffp-contract=off",
            "-fno-fast-math", "-Wno-attributes", "-Wno-unknown-attributes", "-Wno-unknown-pragmas",
            "-U_FORTIFY_SOURCE",
            *SANITIZERS[sanitizer][1]]


def build(variant: str = "base", opt: int = 1, verbose: bool = False,
          sanitizer: str = "asan") -> str:
    cxx = SANITIZERS[sanitizer][0]
    key = hashlib.sha256()
    for p in [os.path.join(CSRC, n) for n in LIB_SOURCES] + \
             [os.path.join(HERE, n) for n in ("simt.hpp", "simt.cpp", "driver.cpp", "build_simt.py")] + \
             [os.path.join(INCLUDE, "redi_hip.h"), os.path.join(HERE, "hip", "hip_runtime.h")]:
        with open(p, "rb") as f:
            key.update(f.read())
    key.update(" ".join(flags(opt, sanitizer) + [variant, cxx]).encode())
    tag = key.hexdigest()[:16]
    out = os.path.join(BUILD, f"redi_simt_{variant}_O{opt}_{sanitizer}")
    stamp = out + ".key"
    if os.path.exists(out) and os.path.exists(stamp) and open(stamp).read() == tag:
        return out
    d = os.path.join(BUILD, f"{variant}_O{opt}_{sanitizer}")
    if os.path.isdir(d):
        shutil.rmtree(d)
    src = os.path.join(d, "src")
    prepare(variant, src)
    inc = ["-I" + HERE, "-I" + INCLUDE, "-I" + src]
    units = [(os.path.join(src, "redi_lib.cpp"), os.path.join(d, "redi_lib.o")),
             (os.path.join(HERE, "simt.cpp"), os.path.join(d, "simt.o")),
             (os.path.join(HERE, "driver.cpp"), os.path.join(d, "driver.o"))]

    def cc(u):
        cmd = [cxx, *flags(opt, sanitizer), *inc, "-c", u[0], "-o", u[1]]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0 or verbose:
            sys.stderr.write(r.stderr[-20000:])
        return r.returncode

    with ThreadPoolExecutor(3) as ex:
        if any(ex.map(cc, units)):
            raise RuntimeError(f"host build ({variant}, -O{opt}, {sanitizer}) failed")
    r = subprocess.run([cxx, *flags(opt, sanitizer), *[u[1] for u in units], "-o", out],
                       capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stderr[-20000:])
        raise RuntimeError("host link failed")
    with open(stamp, "w") as f:
        f.write(tag)
    return out


def build_selftest(sanitizer: str = "asan", opt: int = 1) -> str:
    """tests/simt/selftest.cpp over the same execution model."""
    cxx = SANITIZERS[sanitizer][0]
    srcs = [os.path.join(HERE, n) for n in ("selftest.cpp", "simt.cpp", "simt.hpp")]
    key = hashlib.sha256()
    for p in srcs:
        with open(p, "rb") as f:
            key.update(f.read())
    key.update(" ".join(flags(opt, sanitizer) + [cxx]).encode())
    tag = key.hexdigest()[:16]
    os.makedirs(BUILD, exist_ok=True)
    out = os.path.join(BUILD, f"simt_selftest_O{opt}_{sanitizer}")
    stamp = out + ".key"
    if os.path.exists(out) and os.path.exists(stamp) and open(stamp).read() == tag:
        return out
    src = os.path.join(BUILD, f"selftest_{sanitizer}.cpp")
    with open(srcs[0]) as f:
        text, _ = rewrite(f.read())
    with open(src, "w") as f:
        f.write(f'#line 1 "{srcs[0]}"\n' + text)
    r = subprocess.run([cxx, *flags(opt, sanitizer), "-I" + HERE, src, srcs[1], "-o", out],
                       capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stderr[-20000:])
        raise RuntimeError("self-test build failed")
    with open(stamp, "w") as f:
        f.write(tag)
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--variant", default="base", choices=["base", "nt1"])
    ap.add_argument("--opt", type=int, default=1, choices=[0, 1, 2])
    ap.add_argument("--sanitizer", default="asan", choices=sorted(SANITIZERS))
    ap.add_argument("-v", "--verbose", action="store_true")
    a = ap.parse_args()
    print(build(a.variant, a.opt, a.verbose, a.sanitizer))
    return 0


if __name__ == "__main__":
    sys.exit(main())
