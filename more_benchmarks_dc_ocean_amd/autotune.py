"""Launch-shape autotuner for the Redi kernels.

Single-domain tunables: the thread tile (tx x ty lanes, 64..256) of the
sweep kernel, the tracer group (tracers per flux launch, 1..4), the storage
order of the per-step reconstruction (layer-major or per column) and its
layer chunk (layers per tiled-reconstruction workgroup, 8/16/32).  A tile of
adjacent faces shares columns in cache; a larger tracer group shares one
pass over the face records between more tracers at the price of registers;
a deeper layer chunk trades fewer columns per workgroup (shorter coalesced
row reads) for less halo re-reading.

Batched tunables, per domain of a batch: the 64-lane sweep tile
(tile_i x tile_j) and the reconstruction layer chunk (kchunk) -- the
(tile_i, tile_j, k-chunk) of a domain in the batched kernels, where every
domain carries its own shape in the work table.  The sweep time depends
only on the tile and the reconstruction time only on the chunk, so each is
chosen by timing its own kernel.

Which wins depends on the domain size and layer count, so choices are made
per size bucket:

    bucket = (nz, round(log2(nx * ny)))

``tune`` / ``tune_batch_shape`` time the candidates on a live domain with
the library's per-kernel HIP-event timing and keep the fastest;
``TuneCache`` stores the winners per bucket in a JSON file and answers
lookups for new domains (nearest bucket with the same nz, else the defaults).
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

DEFAULT = {"tx": 64, "ty": 1, "group": 4, "recon_layer_major": 0, "kchunk": 16}
BATCH_DEFAULT = {"tile_i": 64, "tile_j": 1, "kchunk": 16}

TILES = [(64, 1), (16, 4), (8, 8), (128, 1), (64, 2), (32, 4), (256, 1)]
BATCH_TILES = [(64, 1), (32, 2), (16, 4), (8, 8)]  # the batched sweep's 64-lane tiles
KCHUNKS = [8, 16, 32]
GROUPS = [2, 4]
ORDERS = [0, 1]  # reconstruction per column / layer-major


def bucket_of(nx: int, ny: int, nz: int) -> str:
    return f"nz{nz}_a{int(round(math.log2(max(1, nx * ny))))}"


def _time_config(r, dom, tx: int, ty: int, group: int, order: int,
                 steps: int, kchunk: int = 16) -> Tuple[float, Dict]:
    import torch
    r.set_recon_order(order)
    r.set_recon_chunk(kchunk)
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


KEYS = ("tx", "ty", "group", "recon_layer_major", "kchunk")


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
                             "kchunk": 16, "ms_per_step": ms, "kernels": per})
                if log:
                    log(f"recon {'layer' if order else 'column'}-major tile {tx:3d}x{ty:<3d} "
                        f"group {g}: {ms:9.3f} ms/step  "
                        + " ".join(f"{k}={v:.2f}" for k, v in sorted(per.items())))
    best = min(rows, key=lambda d: d["ms_per_step"])
    if not best["recon_layer_major"]:  # the chunk applies to the per-column tiled form
        for kc in KCHUNKS:
            if kc == 16:
                continue
            ms, per = _time_config(r, dom, best["tx"], best["ty"], best["group"], 0, steps, kc)
            rows.append({**{k: best[k] for k in KEYS}, "kchunk": kc, "ms_per_step": ms,
                         "kernels": per})
            if log:
                log(f"  layer chunk {kc:2d}: {ms:9.3f} ms/step")
        best = min(rows, key=lambda d: d["ms_per_step"])
    r.close()
    return {"bucket": bucket_of(case.nx, case.ny, case.nz), "best": {k: best[k] for k in KEYS},
            "best_ms_per_step": best["ms_per_step"], "table": rows,
            "domain": [case.nx, case.ny, case.nz, case.nt]}


def tune_batch_shape(case, form: str = "fp32_store", device: str = "cuda:0",
                     tiles: Iterable = BATCH_TILES, kchunks: Iterable = KCHUNKS,
                     steps: int = 3, log=None) -> Dict:
    """The batched kernels' per-domain shape for ``case``'s bucket: the
    64-lane sweep tile with the least sweep time (u + v grids) and the layer
    chunk with the least reconstruction time, each timed on the
    single-domain kernels, which run the same bodies."""
    from .redi import make_redi
    r, dom = make_redi(case, device, form)
    rows: List[Dict] = []
    for (ti, tj) in tiles:
        for kc in kchunks:
            _, per = _time_config(r, dom, ti, tj, 4, 0, steps, kc)
            rows.append({"tile_i": ti, "tile_j": tj, "kchunk": kc,
                         "sweep_ms": per.get("k_neutral_sweep_u", 0.0) + per.get("k_neutral_sweep_v", 0.0),
                         "recon_ms": per.get("k_tracer_recon", 0.0)})
            if log:
                log(f"tile {ti:2d}x{tj:<2d} chunk {kc:2d}: sweep {rows[-1]['sweep_ms']:.4f} ms, "
                    f"recon {rows[-1]['recon_ms']:.4f} ms")
    r.close()
    # each kernel's time averaged over the other knob
    def avg(key, knob, val):
        v = [x[key] for x in rows if x[knob] == val]
        return sum(v) / len(v)
    bt = min({(x["tile_i"], x["tile_j"]) for x in rows},
             key=lambda t: sum(x["sweep_ms"] for x in rows if (x["tile_i"], x["tile_j"]) == t))
    bk = min({x["kchunk"] for x in rows}, key=lambda k: avg("recon_ms", "kchunk", k))
    best = {"tile_i": bt[0], "tile_j": bt[1], "kchunk": bk}
    dflt = [x for x in rows if (x["tile_i"], x["tile_j"], x["kchunk"]) ==
            (BATCH_DEFAULT["tile_i"], BATCH_DEFAULT["tile_j"], BATCH_DEFAULT["kchunk"])]
    won = [x for x in rows if (x["tile_i"], x["tile_j"], x["kchunk"]) == (bt[0], bt[1], bk)]
    return {"bucket": f"batch_{form}_" + bucket_of(case.nx, case.ny, case.nz), "best": best,
            "best_ms_per_step": (won[0]["sweep_ms"] + won[0]["recon_ms"]) if won else None,
            "default_ms": (dflt[0]["sweep_ms"] + dflt[0]["recon_ms"]) if dflt else None,
            "table": rows, "domain": [case.nx, case.ny, case.nz, case.nt]}


class TuneCache:
    """Per-bucket launch shapes backed by a JSON file."""

    def __init__(self, path: Optional[str] = None):
        self.path = path
        self.entries: Dict[str, Dict] = {}
        if path and os.path.exists(path):
            with open(path) as f:
                self.entries = json.load(f)

    def put(self, result: Dict) -> None:
        e = {"best": result["best"], "ms_per_step": result["best_ms_per_step"],
             "domain": result["domain"]}
        if result.get("default_ms") is not None:
            e["default_ms"] = result["default_ms"]
        self.entries[result["bucket"]] = e

    def save(self) -> None:
        if self.path:
            with open(self.path, "w") as f:
                json.dump(self.entries, f, indent=1, sort_keys=True)

    def _nearest(self, prefix: str, nx: int, ny: int, nz: int) -> Optional[Dict]:
        key = prefix + bucket_of(nx, ny, nz)
        if key in self.entries:
            return dict(self.entries[key]["best"])
        same_nz = [(k, v) for k, v in self.entries.items() if k.startswith(f"{prefix}nz{nz}_")]
        if same_nz:
            a = int(key.split("_a")[1])
            k, v = min(same_nz, key=lambda kv: abs(int(kv[0].split("_a")[1]) - a))
            return dict(v["best"])
        return None

    def lookup(self, nx: int, ny: int, nz: int) -> Dict:
        return self._nearest("", nx, ny, nz) or dict(DEFAULT)

    def batch_shape(self, nx: int, ny: int, nz: int, form: str = "fp32_store") -> Dict:
        """(tile_i, tile_j, kchunk) of a domain in a batch of this form."""
        return self._nearest(f"batch_{form}_", nx, ny, nz) or dict(BATCH_DEFAULT)

    def apply(self, r, nx: int, ny: int, nz: int) -> Dict:
        """Set the bucket's launch shape on an OceanRedi."""
        cfg = dict(DEFAULT)
        cfg.update({k: v for k, v in self.lookup(nx, ny, nz).items() if k in KEYS})
        r.set_recon_order(cfg["recon_layer_major"])
        r.set_recon_chunk(cfg["kchunk"])
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
    ap.add_argument("--batch-form", default="", help="tune the batched per-domain shape "
                    "(tile_i, tile_j, kchunk) of this form for the --sides x --nz buckets")
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
        if args.batch_form:
            res = tune_batch_shape(c, form=args.batch_form, steps=args.steps, log=log)
        else:
            res = tune(c, steps=args.steps, log=log)
        log(f"best {res['best']} {res['best_ms_per_step']:.3f} ms/step ({time.time() - t0:.1f}s)")
        cache.put(res)
    cache.save()
    print(json.dumps(cache.entries, sort_keys=True))
    return 0


if __name__ == "__main__":
    sys.exit(main())
