"""Launch-shape autotuner for the Redi kernels.

The tunables are the thread tile (tx x ty lanes, 64..256) of the sweep
kernel, the tracer group (tracers per flux launch, 1..4) and the storage
order of the per-step reconstruction (layer-major or per column).  A tile
of adjacent faces / cells shares columns and v-faces in cache; a larger
tracer group shares one pass over the face records between more tracers at
the price of registers (and so of occupancy).  Which wins depends on the
domain size and layer count, so choices are made per size bucket:

    bucket = (nz, round(log2(nx * ny)))

``tune`` times every candidate on a live domain with the library's
per-kernel HIP-event timing and keeps the fastest; ``TuneCache`` stores the
winners per bucket in a JSON file and answers lookups for new domains
(nearest bucket with the same nz, else the defaults).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from typing import Dict, Iterable, List, Optional, Tuple

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# the committed winners, measured on MI355X
TUNED = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuned_shapes.json")

DEFAULT = {"tx": 64, "ty": 1, "group": 4, "recon_layer_major": 0}

TILES = [(64, 1), (16, 4), (8, 8), (128, 1), (64, 2), (32, 4), (256, 1)]
GROUPS = [2, 4]
ORDERS = [0, 1]  # reconstruction per column / layer-major


def bucket_of(nx: int, ny: int, nz: int) -> str:
    return f"nz{nz}_a{int(round(math.log2(max(1, nx * ny))))}"


def _time_config(r, dom, tx: int, ty: int, group: int, order: int,
                 steps: int) -> Tuple[float, Dict]:
    import torch
    r.set_recon_order(order)
    r.set_tiles(tx, ty)
    r.set_tracer_group(group)
    r.step(dom, 1)  # warm the new shape
    torch.cuda.synchronize()
    # whole steps between two events on the launch stream, then the same
    # steps again with every kernel bracketed for the per-kernel split
    stream = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(steps):
        r.step(dom, 1)
    e1.record(stream)
    torch.cuda.synchronize()
    total = e0.elapsed_time(e1) / steps
    r.timing(True)
    for _ in range(steps):
        r.step(dom, 1)
    kt = r.collect_timing()
    r.timing(False)
    return total, {k: ms / steps for k, (ms, _) in kt.items()}


KEYS = ("tx", "ty", "group", "recon_layer_major")


def tune(case, device: str = "cuda:0", tiles: Iterable = TILES, groups: Iterable = GROUPS,
         orders: Iterable = ORDERS, steps: int = 2, log=None) -> Dict:
    """Time every (reconstruction order, tile, group) on ``case``; returns
    the winner and the table."""
    from .redi import make_redi
    r, dom = make_redi(case, device)
    rows: List[Dict] = []
    for order in orders:
        for (tx, ty) in tiles:
            for g in groups:
                if g > case.nt and g != min(groups):
                    continue
                ms, per = _time_config(r, dom, tx, ty, g, order, steps)
                rows.append({"tx": tx, "ty": ty, "group": g, "recon_layer_major": order,
                             "ms_per_step": ms, "kernels": per})
                if log:
                    log(f"recon {'layer' if order else 'column'}-major tile {tx:3d}x{ty:<3d} "
                        f"group {g}: {ms:9.3f} ms/step  "
                        + " ".join(f"{k}={v:.2f}" for k, v in sorted(per.items())))
    r.close()
    best = min(rows, key=lambda d: d["ms_per_step"])
    return {"bucket": bucket_of(case.nx, case.ny, case.nz), "best": {k: best[k] for k in KEYS},
            "best_ms_per_step": best["ms_per_step"], "table": rows,
            "domain": [case.nx, case.ny, case.nz, case.nt]}


class TuneCache:
    """Per-bucket launch shapes backed by a JSON file."""

    def __init__(self, path: Optional[str] = None):
        self.path = path
        self.entries: Dict[str, Dict] = {}
        if path and os.path.exists(path):
            with open(path) as f:
                self.entries = json.load(f)

    def put(self, result: Dict) -> None:
        self.entries[result["bucket"]] = {"best": result["best"],
                                          "ms_per_step": result["best_ms_per_step"],
                                          "domain": result["domain"]}

    def save(self) -> None:
        if self.path:
            with open(self.path, "w") as f:
                json.dump(self.entries, f, indent=1, sort_keys=True)

    def lookup(self, nx: int, ny: int, nz: int) -> Dict:
        key = bucket_of(nx, ny, nz)
        if key in self.entries:
            return dict(self.entries[key]["best"])
        same_nz = [(k, v) for k, v in self.entries.items() if k.startswith(f"nz{nz}_")]
        if same_nz:
            a = int(key.split("_a")[1])
            k, v = min(same_nz, key=lambda kv: abs(int(kv[0].split("_a")[1]) - a))
            return dict(v["best"])
        return dict(DEFAULT)

    def apply(self, r, nx: int, ny: int, nz: int) -> Dict:
        """Set the bucket's launch shape on an OceanRedi."""
        cfg = dict(DEFAULT)
        cfg.update({k: v for k, v in self.lookup(nx, ny, nz).items() if k in KEYS})
        r.set_recon_order(cfg["recon_layer_major"])
        r.set_tiles(cfg["tx"], cfg["ty"])
        r.set_tracer_group(cfg["group"])
        return cfg


def main() -> int:
    sys.path.insert(0, ROOT)
    from more_benchmarks_dc_ocean_amd import inputs
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="two_fronts_1440x1080x75")
    ap.add_argument("--sides", default="", help="comma list of square interior sides to tune "
                    "(batched buckets), e.g. 48,96,192,384")
    ap.add_argument("--nz", default="75")
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--out", default="")
    args = ap.parse_args()
    log = lambda m: print(f"[autotune] {m}", flush=True)
    cache = TuneCache(args.out or None)
    cases = []
    if args.sides:
        for side in (int(x) for x in args.sides.split(",")):
            for nz in (int(x) for x in args.nz.split(",")):
                cases.append(inputs.case_domain(3, side * 1000 + nz, side, side, nz))
    else:
        builder = {"seamount_64x64x25": inputs.case_seamount, "front_473x297x75": inputs.case_front,
                   "two_fronts_1440x1080x75": inputs.case_two_fronts}[args.workload]
        cases.append(builder(nsteps=1))
    for c in cases:
        log(f"domain {c.nx}x{c.ny}x{c.nz} nt={c.nt} ({bucket_of(c.nx, c.ny, c.nz)})")
        t0 = time.time()
        res = tune(c, steps=args.steps, log=log)
        log(f"best {res['best']} {res['best_ms_per_step']:.3f} ms/step ({time.time() - t0:.1f}s)")
        cache.put(res)
    cache.save()
    print(json.dumps(cache.entries, sort_keys=True))
    return 0


if __name__ == "__main__":
    sys.exit(main())
