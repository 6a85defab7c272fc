"""Workload driver for rocprofv3 runs.

Runs, in one process on cuda:0:
  1. a calibration copy of exactly ``--copy-mib`` MiB (torch ``copy_`` of a
     contiguous float64 tensor: that many MiB read and the same written), so
     FETCH_SIZE / WRITE_SIZE counters can be scaled against known bytes;
  2. ``--warmup`` + ``--steps`` Redi steps of the chosen workload.

Typical use (each counter in its own run):
    rocprofv3 --kernel-trace --stats -d gpurun_out/prof -o run -- \
        python -m more_benchmarks_dc_ocean_amd.profiling --workload two_fronts_1440x1080x75 --steps 3
    rocprofv3 --pmc FETCH_SIZE -d gpurun_out/pmc_fetch -o run -- \
        python -m more_benchmarks_dc_ocean_amd.profiling --workload two_fronts_1440x1080x75 --steps 1
``summarize`` turns the CSV outputs into the per-kernel figures kept under
profiles/.
"""
from __future__ import annotations

import argparse
import csv
import glob
import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def run(workload: str, steps: int, warmup: int, copy_mib: int, form: str = "fp64") -> None:
    import torch
    from more_benchmarks_dc_ocean_amd import inputs
    from more_benchmarks_dc_ocean_amd.autotune import TUNED, TuneCache
    from more_benchmarks_dc_ocean_amd.redi import make_redi
    dev = "cuda:0"
    n = copy_mib * (1 << 20) // 8
    a = torch.ones(n, dtype=torch.float64, device=dev)
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    b.copy_(a)
    torch.cuda.synchronize()
    del a, b
    builder = {"seamount_64x64x25": inputs.case_seamount, "front_473x297x75": inputs.case_front,
               "two_fronts_1440x1080x75": inputs.case_two_fronts}[workload]
    case = builder(nsteps=steps)
    print(f"[profiling] {workload} {case.nx}x{case.ny}x{case.nz} nt={case.nt}", flush=True)
    r, dom = make_redi(case, dev)
    TuneCache(TUNED).apply(r, case.nx, case.ny, case.nz)
    for _ in range(warmup + steps):
        r.step(dom, 1)
    torch.cuda.synchronize()
    print("[profiling] done", flush=True)


def _rows(path_glob: str):
    out = []
    for p in sorted(glob.glob(path_glob, recursive=True)):
        with open(p, newline="") as f:
            out.extend(csv.DictReader(f))
    return out


def summarize_counters(fetch_glob: str, write_glob: str, copy_mib: int) -> dict:
    """Per-kernel average FETCH_SIZE / WRITE_SIZE per dispatch (KiB), scaled by
    the calibration copy (whose true traffic is copy_mib MiB each way)."""
    res = {}
    for name, g in (("FETCH_SIZE", fetch_glob), ("WRITE_SIZE", write_glob)):
        acc = {}
        for row in _rows(g):
            if row.get("Counter_Name", name) != name:
                continue
            k = row.get("Kernel_Name", "")
            v = float(row.get("Counter_Value", 0.0))
            acc.setdefault(k, []).append(v)
        res[name] = {k: sum(v) / len(v) for k, v in acc.items()}
    copy_k = [k for k in res["FETCH_SIZE"] if "copy" in k.lower() or "elementwise" in k.lower()]
    scale = {}
    for name in res:
        ck = [k for k in res[name] if k in copy_k]
        if ck:
            meas = res[name][ck[0]]
            scale[name] = (copy_mib * 1024.0) / meas if meas > 0 else None
        else:
            scale[name] = None
    return {"raw_KiB_per_dispatch": res, "scale": scale, "copy_kernels": copy_k}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="two_fronts_1440x1080x75")
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--copy-mib", type=int, default=2048)
    ap.add_argument("--summarize", nargs=2, metavar=("FETCH_GLOB", "WRITE_GLOB"))
    args = ap.parse_args()
    if args.summarize:
        print(json.dumps(summarize_counters(args.summarize[0], args.summarize[1], args.copy_mib),
                         indent=1))
        return 0
    run(args.workload, args.steps, args.warmup, args.copy_mib)
    return 0


if __name__ == "__main__":
    sys.exit(main())
