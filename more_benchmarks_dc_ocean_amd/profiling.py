"""Workload driver for rocprofv3 runs.

Runs, in one process on cuda:0:
  1. a calibration copy of exactly ``--copy-mib`` MiB (torch ``copy_`` of a
     contiguous float64 tensor: that many MiB read and the same written), so
     FETCH_SIZE / WRITE_SIZE counters can be scaled against known bytes;
  2. ``--warmup`` + ``--steps`` Redi steps of the chosen workload.

Typical use (each counter in its own run):
    rocprofv3 --kernel-trace --stats -d prof_out/kt -o run -- \
        python -m more_benchmarks_dc_ocean_amd.profiling --workload two_fronts_1440x1080x75 --steps 3
    rocprofv3 --pmc FETCH_SIZE -d prof_out/fetch -o run -- \
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


BATCHED = "batched_96"  # the batched workload: 96 uneven domains (inputs.case_batch)


def _cases(workload: str, steps: int):
    from more_benchmarks_dc_ocean_amd import inputs
    if workload == BATCHED:
        return inputs.case_batch(96, seed=0, nsteps=steps)
    builder = {"seamount_64x64x25": inputs.case_seamount, "front_473x297x75": inputs.case_front,
               "two_fronts_1440x1080x75": inputs.case_two_fronts,
               "gabight_473x297x75": inputs.case_gabight,
               "analytic_473x297x30": inputs.case_analytic}[workload]
    return [builder(nsteps=steps)]


def run(workload: str, steps: int, warmup: int, copy_mib: int, form: str = "fp64") -> None:
    import torch
    from more_benchmarks_dc_ocean_amd.autotune import TUNED, TuneCache
    from more_benchmarks_dc_ocean_amd.redi import DeviceDomain, make_redi
    dev = "cuda:0"
    n = copy_mib * (1 << 20) // 8
    a = torch.ones(n, dtype=torch.float64, device=dev)
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    b.copy_(a)
    torch.cuda.synchronize()
    del a, b
    cases = _cases(workload, steps)
    cache = TuneCache(TUNED)
    if workload == BATCHED:
        from more_benchmarks_dc_ocean_amd.batch import RediBatch, domain_bytes
        arena = sum(domain_bytes(c.nx, c.ny, c.nz, form, 4) for c in cases)
        bt = RediBatch(form=form, group=4, max_domains=len(cases), arena_bytes=arena, device=dev)
        doms = []
        for c in cases:
            doms.append(DeviceDomain(c, dev, form))
            bt.add(doms[-1], cache.batch_shape(c.nx, c.ny, c.nz, form))
        print(f"[profiling] {workload} {len(cases)} domains, form {form}", flush=True)
        bt.step(warmup + steps)
        torch.cuda.synchronize()
        bt.close()
        print("[profiling] done", flush=True)
        return
    case = cases[0]
    print(f"[profiling] {workload} {case.nx}x{case.ny}x{case.nz} nt={case.nt} form {form}", flush=True)
    r, dom = make_redi(case, dev, form)
    cache.apply(r, case.nx, case.ny, case.nz)
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


# the device-to-device copy torch issues for the calibration tensor
CALIBRATION_KERNEL = "__amd_rocclr_copyBuffer"
GPU_CLOCK_HZ = 2.4e9   # MI355X peak engine clock (agent info: 2400 MHz)
N_CU, N_SIMD = 256, 1024


def _short(name: str) -> str:
    return name.split("(")[0].replace("void ", "").replace("redi::", "")


def _base(name: str) -> str:
    """The byte model's name for a kernel (template arguments, the
    reconstruction's tiled variant and the batched kernels' kb_ prefix folded
    in)."""
    short = _short(name)
    b = short.split("<")[0]
    if b.startswith("kb_"):
        b = "k_" + b[3:]
    if b == "k_flux":  # k_flux<F, NT, NZC, y>: the x (u faces) or y (v faces) grid
        b = "k_flux_y" if short.rstrip(">").split(",")[-1].strip() == "true" else "k_flux_x"
    return "k_tracer_recon" if b.startswith("k_tracer_recon") else b


def _flux_class(name: str):
    """The layer-capacity class of a k_flux / kb_flux kernel name."""
    return int(_short(name).rstrip(">").split(",")[-2])


def summarize_dir(d: str, workload: str, copy_mib: int = 2048, form: str = "fp64") -> dict:
    """Per-kernel figures from one directory of runs: kt/ (kernel trace
    --stats), fetch/ and write/ (FETCH_SIZE, WRITE_SIZE) and any other
    --pmc passes (SQ_*, TA_*, TD_*, TCC_* counters), all per dispatch."""
    from . import byte_model
    from .autotune import TUNED, TuneCache
    cases = _cases(workload, 1)
    alg: dict = {}
    flux_by_cap: dict = {}  # batched: the flux kernels' bytes per (grid, layer-capacity class)
    for case in cases:
        if workload == BATCHED:
            cap = next(c for c in (32, 76, 128, 256) if c >= case.nz)
            sl = byte_model.step_launches(case, 4, form)
            for kk in ("k_flux_x", "k_flux_y"):
                flux_by_cap[(kk, cap)] = flux_by_cap.get((kk, cap), 0.0) + sum(sl[kk])
        group = 4 if workload == BATCHED else \
            TuneCache(TUNED).lookup(case.nx, case.ny, case.nz).get("group", 4)
        for k, v in byte_model.step_launches(case, group, form).items():
            if not v:
                continue
            if workload == BATCHED:  # one launch per kind covers every domain (u and v together)
                kk = "k_neutral_sweep" if k.startswith("k_neutral_sweep") else k
                alg[kk] = alg.get(kk, 0.0) + sum(v)
            else:
                alg[k] = sum(v) / len(v)
    if workload == BATCHED:
        alg["k_neutral_sweep_u"] = alg.pop("k_neutral_sweep", 0.0)
    stats = {}
    for row in _rows(os.path.join(d, "kt", "*kernel_stats.csv")):
        stats[_short(row["Name"])] = (int(row["Calls"]), float(row["AverageNs"]) * 1e-6)
    counters: dict = {}
    for row in _rows(os.path.join(d, "*", "*counter_collection.csv")):
        k = _short(row["Kernel_Name"])
        c = counters.setdefault(k, {}).setdefault(row["Counter_Name"], [])
        c.append(float(row["Counter_Value"]))
    per = {k: {c: sum(v) / len(v) for c, v in cs.items()} for k, cs in counters.items()}
    cal = per.get(CALIBRATION_KERNEL, {})
    fscale = copy_mib * 1024.0 / cal["FETCH_SIZE"] if cal.get("FETCH_SIZE") else None
    wscale = copy_mib * 1024.0 / cal["WRITE_SIZE"] if cal.get("WRITE_SIZE") else None
    out = {"workload": workload, "form": form, "fetch_scale": fscale, "write_scale": wscale,
           "kernels": {}}
    for k, (calls, ms) in sorted(stats.items(), key=lambda kv: -kv[1][0] * kv[1][1]):
        if not k.startswith(("k_", "kb_")):
            continue
        c = per.get(k, {})
        cyc = ms * 1e-3 * GPU_CLOCK_HZ
        e = {"calls": calls, "avg_ms": ms}
        a = alg.get(_base(k) if _base(k) != "k_neutral_sweep" else "k_neutral_sweep_u")
        if workload == BATCHED and k.startswith("kb_flux"):
            a = flux_by_cap.get((_base(k), _flux_class(k)))
        if a is not None:
            e["alg_bytes"] = a
            e["alg_GBps"] = a / (ms * 1e-3) / 1e9
        if "FETCH_SIZE" in c and fscale:
            e["fetch_bytes"] = c["FETCH_SIZE"] * 1024.0 * fscale
        if "WRITE_SIZE" in c and wscale:
            e["write_bytes"] = c["WRITE_SIZE"] * 1024.0 * wscale
        if "fetch_bytes" in e and "write_bytes" in e:
            e["traffic_bytes"] = e["fetch_bytes"] + e["write_bytes"]
            e["traffic_GBps"] = e["traffic_bytes"] / (ms * 1e-3) / 1e9
        if "SQ_ACTIVE_INST_VALU" in c:
            e["valu_busy"] = c["SQ_ACTIVE_INST_VALU"] * 4.0 / N_SIMD / cyc
        if "TA_BUSY_avr" in c:
            e["ta_busy"] = c["TA_BUSY_avr"] / cyc
        if "TD_TD_BUSY_sum" in c:
            e["td_busy"] = c["TD_TD_BUSY_sum"] / N_CU / cyc
        if "TCC_HIT_sum" in c and "TCC_MISS_sum" in c:
            tot = c["TCC_HIT_sum"] + c["TCC_MISS_sum"]
            e["l2_hit"] = c["TCC_HIT_sum"] / tot if tot else None
        for n in ("SQ_WAVES", "SQ_INSTS_VALU", "SQ_INSTS_VMEM_RD"):
            if n in c:
                e[n] = c[n]
        out["kernels"][k] = e
    return out


def markdown(summary: dict, title: str, commands: list) -> str:
    f = lambda v, n=2: "" if v is None else f"{v:.{n}f}"
    pct = lambda v: "" if v is None else f"{100 * v:.0f}%"
    L = [f"# {title}", "", "Commands (repository root, one MI355X, `export TMPDIR=/tmp`):", "", "```"]
    L += commands + ["```", ""]
    L.append(f"Calibration: the driver's {2048} MiB device copy (`{CALIBRATION_KERNEL}`) reads "
             f"FETCH_SIZE x {f(summary['fetch_scale'], 5)} and WRITE_SIZE x "
             f"{f(summary['write_scale'], 5)} of the bytes it moves; the traffic columns are "
             "scaled by these factors.  `alg` is the kernel's share of the algorithmic bytes of "
             "one step (compulsory traffic of the two-call contract), per launch.  Busy "
             "fractions are per dispatch over the kernel's average duration at 2.4 GHz.")
    L += ["", "| kernel | calls | avg ms | alg GB | alg GB/s | fetch GB | write GB | traffic GB/s | "
          "VALU busy | TA busy | TD busy | L2 hit |", "|---|---|---|---|---|---|---|---|---|---|---|---|"]
    for k, e in summary["kernels"].items():
        g = lambda n: e.get(n)
        L.append(f"| `{k}` | {e['calls']} | {f(e['avg_ms'], 3)} | "
                 f"{f(g('alg_bytes') / 1e9 if g('alg_bytes') is not None else None)} | "
                 f"{f(g('alg_GBps'), 0)} | "
                 f"{f(g('fetch_bytes') / 1e9 if g('fetch_bytes') is not None else None)} | "
                 f"{f(g('write_bytes') / 1e9 if g('write_bytes') is not None else None)} | "
                 f"{f(g('traffic_GBps'), 0)} | {pct(g('valu_busy'))} | {pct(g('ta_busy'))} | "
                 f"{pct(g('td_busy'))} | {pct(g('l2_hit'))} |")
    return "\n".join(L) + "\n"


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
    copy_k = [k for k in res["FETCH_SIZE"] if k.startswith(CALIBRATION_KERNEL)]
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
    ap.add_argument("--summarize-dir", metavar="DIR",
                    help="directory with kt/, fetch/, write/ and other --pmc pass outputs")
    ap.add_argument("--json", metavar="FILE", help="with --summarize-dir: also write the figures")
    ap.add_argument("--markdown", metavar="FILE", help="with --summarize-dir: also write the table")
    ap.add_argument("--title", default="")
    ap.add_argument("--commands", default="", help="file with the commands that made the runs")
    ap.add_argument("--form", default="fp64", choices=["fp64", "fp32_store"])
    args = ap.parse_args()
    if args.summarize_dir:
        summ = summarize_dir(args.summarize_dir, args.workload, args.copy_mib, args.form)
        if args.json:
            with open(args.json, "w") as fh:
                json.dump(summ, fh, indent=1)
        if args.markdown:
            cmds = open(args.commands).read().strip().splitlines() if args.commands else []
            with open(args.markdown, "w") as fh:
                fh.write(markdown(summ, args.title or f"{args.workload} {args.form}", cmds))
        print(json.dumps(summ, indent=1))
        return 0
    if args.summarize:
        print(json.dumps(summarize_counters(args.summarize[0], args.summarize[1], args.copy_mib),
                         indent=1))
        return 0
    run(args.workload, args.steps, args.warmup, args.copy_mib, args.form)
    return 0


if __name__ == "__main__":
    sys.exit(main())
