"""Redi step benchmark on one MI355X.

Workload (default): the 1440x1080x75 domain, 4 tracers, Roquet SpV EOS,
~30% land, two fronts plus mesoscale noise, dt = 900 s, kappa = 1000 m^2/s,
all-wall edges (nghost = 3, so arrays are 1446x1086x75).  One "step" is the
reference's timed region: redi_calc_coeffs (Phase A) followed by
redi_apply_flux over all four tracers (Phase B), every tracer updated in
place.  Inputs are generated on the host, copied to HBM, and resident there
before the timed region starts.

Both precision forms are timed on the same inputs, one after the other:
the plain FP64 form gives ``value``; the FP32-store form (fields, face
records and fluxes stored in FP32) is reported under ``forms.fp32_store``
with its own roofline, its ratio to the FP64 form and ``accuracy``: its
error against the FP64 form after ``--accuracy-steps`` steps and how many
down-gradient guard decisions it switches, beside what rounding the inputs
to FP32 alone does to the FP64 form.  ``--form`` times one
form only.

Printed JSON line: metric "Redi Mcell-steps/s" (interior cells x steps per
second), plus
  roofline      the dominant kernel's algorithmic bytes per launch divided by
                its average launch time (HIP events around every launch on
                the launch stream, recorded inside the timed region), against
                the 8 TB/s HBM peak; ``traffic`` is the rocprofv3-measured
                HBM bytes per launch when a committed measurement for this
                kernel and workload exists under profiles/;
  cpu_baseline  the reference's own Fortran Redi (oracle/_ref/redi_ref, built
                from the reference sources) timed on this host's cores over
                a bounded sample of the same workload, or the oracle's C
                restatement when that binary is absent.

  real_bathymetry
                both forms on the reference's real 0.1-degree gabight
                bathymetry (473x297x75, 4 tracers).
  reference_benchmark
                the FP64 form on the reference's own benchmark state
                (473x297x30, all wet, linear EOS, 2 tracers) beside the
                reference's published V100 times for that configuration.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload two_fronts_1440x1080x75]
                       [--form both|fp64|fp32_store] [--reference-steps 20]
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0

WORKLOADS = {
    # name: (builder, description)
    "two_fronts_1440x1080x75": ("case_two_fronts", "redi_1440x1080x75_roquet_4tr"),
    "front_473x297x75": ("case_front", "redi_473x297x75_roquet_4tr"),
    "seamount_64x64x25": ("case_seamount", "redi_64x64x25_linear_2tr"),
}


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def measured_traffic(kernel: str, workload: str, form: str = "fp64"):
    """HBM bytes per launch (copy-calibrated FETCH_SIZE + WRITE_SIZE) of
    ``kernel`` from the committed rocprofv3 summary of this workload."""
    p = os.path.join(ROOT, "profiles", f"kernels_{workload}_{form}.json")
    if not os.path.exists(p):
        return None
    with open(p) as f:
        d = json.load(f)
    from more_benchmarks_dc_ocean_amd.profiling import _base
    for k, e in d.get("kernels", {}).items():
        if _base(k) == kernel and "traffic_bytes" in e:
            return e["traffic_bytes"]
    return None


def _cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _time_reference(case, steps: int, threads: int) -> float:
    """ms per step of the reference build on ``case`` (1 warm-up step)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import redi_oracle
    _, out = redi_oracle.run_reference(case, nsteps=steps, threads=threads, nwarm=1,
                                       want_output=False)
    return float(out.split("ms_per_step=")[1].split()[0])


def cpu_baseline(case, sample_steps: int = 2, extra: bool = True):
    """The reference's own Fortran build (oracle/_ref/redi_ref) on this
    host's cores.  The headline figure: a 480x360 interior window of the
    workload domain, all layers and tracers, 1 warm-up + ``sample_steps``
    timed steps on every core.  With ``extra``, also the CPU-runnable
    64x64x25 configuration (10 steps) and the 473x297x75 front (1 step), each
    on 1 thread and on every core, and one step of the whole workload domain
    (after 1 warm-up) on every core.  Falls back to the oracle's C restatement
    (kind "port") only where the reference binary is absent."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import redi_oracle
    from more_benchmarks_dc_ocean_amd import inputs
    from more_benchmarks_dc_ocean_amd.inputs import crop
    nxp = min(480, case.nx_phys)
    nyp = min(360, case.ny_phys)
    i0 = (case.nx_phys - nxp) // 2
    j0 = (case.ny_phys - nyp) // 2
    sub = crop(case, i0, j0, nxp, nyp)
    nproc = os.cpu_count() or 1
    threads = min(nproc, 64)
    cells = sub.nx_phys * sub.ny_phys * sub.nz
    sample = (f"{nxp}x{nyp}x{sub.nz} interior window of the workload domain, {sub.nt} tracers, "
              f"{sample_steps} timed steps after 1 warm-up")
    prov = {"cpu_model": _cpu_model(), "nproc": nproc}
    if redi_oracle.have_reference():
        try:
            ms = _time_reference(sub, sample_steps, threads)
            res = {"value": cells / (ms * 1e-3) / 1e6, "unit": "Mcell-steps/s", "cores": threads,
                   "kind": "reference", "sample": sample, "ms_per_step": ms, **prov}
            if extra:
                runs = []
                for name, c, steps in (("seamount_64x64x25 (linear EOS, 2 tracers)",
                                        inputs.case_seamount(nsteps=10), 10),
                                       ("front_473x297x75 (Roquet, 4 tracers)",
                                        inputs.case_front(nsteps=1), 1)):
                    for th in (1, threads):
                        m = _time_reference(c, steps, th)
                        runs.append({"workload": name, "threads": th, "steps": steps,
                                     "ms_per_step": round(m, 3),
                                     "Mcell_steps_per_s": round(c.interior_cells / (m * 1e-3) / 1e6, 3)})
                        log(f"cpu reference {name} on {th} thread(s): {m:.1f} ms/step")
                m = _time_reference(case, 1, threads)  # the whole workload domain
                runs.append({"workload": f"{case.nx_phys}x{case.ny_phys}x{case.nz} whole domain "
                                         f"({case.nt} tracers)", "threads": threads, "steps": 1,
                             "ms_per_step": round(m, 3),
                             "Mcell_steps_per_s": round(case.interior_cells / (m * 1e-3) / 1e6, 3)})
                log(f"cpu reference whole domain on {threads} threads: {m:.1f} ms/step")
                res["runs"] = runs
            return res
        except Exception as e:  # the binary exists but cannot run here
            log(f"reference CPU build failed ({e}); timing the C restatement instead")
    redi_oracle.set_threads(threads)
    o = redi_oracle.OracleRedi(sub)
    h = np.ascontiguousarray(sub.h)
    hTr = sub.hTr.copy()
    o.step(h, hTr)
    t0 = time.perf_counter()
    for _ in range(sample_steps):
        o.step(h, hTr)
    ms = (time.perf_counter() - t0) * 1e3 / sample_steps
    return {"value": cells / (ms * 1e-3) / 1e6, "unit": "Mcell-steps/s", "cores": threads,
            "kind": "port", "sample": sample, "ms_per_step": ms, **prov}


def fp32_accuracy(case, steps: int) -> dict:
    """Error of the FP32-store form after ``steps`` steps of the workload,
    against the FP64 form, and the down-gradient guard decisions of step 1
    (more_benchmarks_dc_ocean_amd.accuracy: A = FP64 form on the FP64 inputs,
    B = FP64 form on the inputs rounded to FP32, C = FP32-store form on the
    rounded inputs).  Relative errors are to the reference field's max."""
    from more_benchmarks_dc_ocean_amd import accuracy
    res = accuracy.report(case, steps, log=lambda m: log(f"[accuracy] {m}"))
    F, out = res["fields"], {}
    for n, g in res["guard"].items():
        e = lambda key: F[key]["max_rel"]
        out[n] = {"tend_max_rel_C_vs_A": e(f"tend_{n} (C vs A)"),
                  "tend_rms_rel_C_vs_A": F[f"tend_{n} (C vs A)"]["rms_rel"],
                  "tend_max_rel_B_vs_A": e(f"tend_{n} (B vs A)"),
                  "tend_max_rel_C_vs_B": e(f"tend_{n} (C vs B)"),
                  f"hTr_max_rel_after_{steps}_C_vs_A": e(f"hTr_{n} after {steps} (C vs A)"),
                  "guard_zeroed_A": g["zeroed_A"], "guard_flips_B_vs_A": g["flips_B_vs_A"],
                  "guard_flips_C_vs_A": g["flips_C_vs_A"], "guard_flips_C_vs_B": g["flips_C_vs_B"]}
    for n in ("T", "S"):
        out[n][f"layer_value_max_rel_after_{steps}_C_vs_A"] = \
            F[f"{n} layer value after {steps} (C vs A)"]["max_rel"]
    return {"steps": steps,
            "runs": "A fp64 on fp64 inputs, B fp64 on inputs rounded to fp32, C fp32_store on the "
                    "rounded inputs; relative to the reference field's max",
            "tracers": {n: {k: (float(f"{v:.4g}") if isinstance(v, float) else v) for k, v in d.items()}
                        for n, d in out.items()}}


def time_form(case, form: str, steps: int, warmup: int, workload: str) -> dict:
    """Time ``steps`` Redi steps of ``form`` on resident inputs; the
    dominant kernel's roofline from HIP events around its launches."""
    import torch
    from more_benchmarks_dc_ocean_amd import byte_model
    from more_benchmarks_dc_ocean_amd.autotune import TUNED, TuneCache
    from more_benchmarks_dc_ocean_amd.redi import make_redi

    dev = "cuda:0"
    r, dom = make_redi(case, dev, form)
    shape = TuneCache(TUNED).apply(r, case.nx, case.ny, case.nz)
    log(f"[{form}] launch shape {shape}")
    stream = torch.cuda.current_stream(dev)
    for _ in range(warmup):
        r.step(dom, 1)
    torch.cuda.synchronize()
    log(f"[{form}] warm-up {warmup} steps done")

    r.timing(True)
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    wall0 = time.perf_counter()
    ev0.record(stream)
    for _ in range(steps):
        r.step(dom, 1)
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - wall0
    ms_total = ev0.elapsed_time(ev1)
    kt = r.collect_timing()
    r.timing(False)
    ms_per_step = ms_total / steps
    value = case.interior_cells * steps / (ms_total * 1e-3) / 1e6
    log(f"[{form}] timed {steps} steps: {ms_per_step:.3f} ms/step (wall {wall * 1e3 / steps:.3f})")
    for k, (ms, n) in sorted(kt.items(), key=lambda kv: -kv[1][0]):
        log(f"  {k:20s} {ms / n:9.4f} ms/launch x{n:4d}  ({100 * ms / ms_total:5.1f}% of step time)")

    dom_k = max(kt.items(), key=lambda kv: kv[1][0])[0]
    avg_ms = kt[dom_k][0] / kt[dom_k][1]
    alg = byte_model.launch_bytes(case, dom_k.split("<")[0], shape["group"], form)
    achieved = alg / (avg_ms * 1e-3) / 1e9
    step_gbs = byte_model.step_bytes(case, form) * steps / (ms_total * 1e-3) / 1e9
    roof = {"bound": "hbm", "kernel": dom_k, "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
            "traffic": measured_traffic(dom_k, workload, form),
            "alg_bytes_per_launch": alg, "avg_launch_ms": round(avg_ms, 4),
            "launches_per_step": kt[dom_k][1] // steps,
            "step_alg_bytes_per_cell": round(byte_model.bytes_per_cell_step(case, form), 2),
            "step_alg_GBps": round(step_gbs, 2), "step_frac": round(step_gbs / HBM_PEAK_GBS, 5)}
    kernels = {k: round(ms / n, 4) for k, (ms, n) in kt.items()}
    r.close()
    del dom
    torch.cuda.empty_cache()
    return {"value": round(value, 3), "ms_per_step": round(ms_per_step, 4), "roofline": roof,
            "launch_shape": shape, "kernel_ms_per_launch": kernels}


# The reference's published times for its own Redi benchmark configuration
# (473x297x30, all wet, analytic T/S, linear EOS, 2 tracers, 20 timed calls
# after 10 warm-up calls of redi_calc_coeffs + redi_apply_flux), one V100.
REFERENCE_V100_MS = {
    "DC (production do-concurrent Fortran)": (44.72, "legacy_testing/redi_benchmark/README.md:9"),
    "CUDA faithful port": (44.98, "legacy_testing/redi_benchmark/README.md:10"),
    "DC + PRECOMP": (33.22, "legacy_testing/redi_benchmark/README.md:11"),
    "CUDA optimised": (42.11, "redi/OPTIMIZATION.md:40-42"),
}


def time_reference_config(steps: int, warmup: int = 10) -> dict:
    """The FP64 form on the reference's own benchmark state (inputs.case_analytic,
    473x297x30), timed as the reference times it; beside it the reference's
    published V100 times for the same configuration (other hardware)."""
    from more_benchmarks_dc_ocean_amd import inputs
    case = inputs.case_analytic(nsteps=steps)
    r = time_form(case, "fp64", steps, warmup, case.name)
    best = min(v[0] for v in REFERENCE_V100_MS.values())
    return {"workload": case.name, "nx_phys": case.nx_phys, "ny_phys": case.ny_phys, "nz": case.nz,
            "ntracers": case.nt, "eos": "linear", "form": "fp64", "steps": steps, "warmup": warmup,
            "ms_per_step": r["ms_per_step"], "value": r["value"], "unit": "Mcell-steps/s",
            "published_v100_ms_per_step": {k: {"ms": v[0], "source": v[1]}
                                          for k, v in REFERENCE_V100_MS.items()},
            "speedup_vs_fastest_published": round(best / r["ms_per_step"], 2)}


def time_gabight(steps: int, warmup: int) -> dict:
    """Both forms on the real Great Australian Bight bathymetry at its native
    473x297 resolution, 75 layers, 4 tracers (inputs.case_gabight)."""
    from more_benchmarks_dc_ocean_amd import inputs
    case = inputs.case_gabight(nsteps=steps)
    out = {"workload": case.name, "nx_phys": case.nx_phys, "ny_phys": case.ny_phys, "nz": case.nz,
           "ntracers": case.nt, "eos": "roquet_spv", "land_fraction": round(case.land_fraction, 4),
           "steps": steps, "unit": "Mcell-steps/s"}
    for form in ("fp64", "fp32_store"):
        r = time_form(case, form, steps, warmup, case.name)
        out[form] = {"value": r["value"], "ms_per_step": r["ms_per_step"]}
    return out


def time_batched(cases, form: str, steps: int, warmup: int, persist: int = 1,
                 tuned: bool = True) -> dict:
    """Batched workload: every domain of ``cases`` in one RediBatch, timed
    per step (one hipGraph launch, persistent kernels over all domains) and,
    on the same resident state, as one launch sequence per domain."""
    import torch
    from more_benchmarks_dc_ocean_amd.autotune import TUNED, TuneCache
    from more_benchmarks_dc_ocean_amd.batch import RediBatch, domain_bytes
    from more_benchmarks_dc_ocean_amd.redi import DeviceDomain

    dev = "cuda:0"
    cache = TuneCache(TUNED)
    arena = sum(domain_bytes(c.nx, c.ny, c.nz, form, 4) for c in cases)
    b = RediBatch(form=form, group=4, max_domains=len(cases), arena_bytes=arena, persist=persist,
                  device=dev)
    doms, shapes = [], []
    for c in cases:
        d = DeviceDomain(c, dev, form)
        shape = cache.batch_shape(c.nx, c.ny, c.nz, form) if tuned else None
        b.add(d, shape)
        doms.append(d)
        shapes.append(shape)
    cells = sum(c.interior_cells for c in cases)
    stream = torch.cuda.current_stream(dev)

    def timed(fn) -> float:
        for _ in range(warmup):
            fn(1)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(steps):
            fn(1)
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / steps

    ms_batched = timed(b.step)
    log(f"[batched {form}] {len(cases)} domains, {cells / 1e6:.1f} M cells: {ms_batched:.3f} ms/step")
    ms_serial = timed(b.step_serial)
    log(f"[batched {form}] one launch sequence per domain: {ms_serial:.3f} ms/step")
    from more_benchmarks_dc_ocean_amd import byte_model
    alg = sum(byte_model.step_bytes(c, form) for c in cases)
    b.close()
    del doms
    torch.cuda.empty_cache()
    hist = {}
    for sh in shapes:
        k = "default" if sh is None else f"{sh['tile_i']}x{sh['tile_j']}/k{sh['kchunk']}"
        hist[k] = hist.get(k, 0) + 1
    return {"domains": len(cases), "cells": cells, "form": form, "persist": persist,
            "value": round(cells / (ms_batched * 1e-3) / 1e6, 3), "unit": "Mcell-steps/s",
            "ms_per_step": round(ms_batched, 4),
            "serial_value": round(cells / (ms_serial * 1e-3) / 1e6, 3),
            "serial_ms_per_step": round(ms_serial, 4),
            "speedup_vs_serial": round(ms_serial / ms_batched, 4),
            "step_alg_GBps": round(alg / (ms_batched * 1e-3) / 1e9, 2),
            "step_frac": round(alg / (ms_batched * 1e-3) / 1e9 / HBM_PEAK_GBS, 5),
            "shapes": hist, "arena_GiB": round(arena / 2**30, 2)}


def _copy_requests(reqs):
    from more_benchmarks_dc_ocean_amd import serve
    return [serve.Request(rid=r.rid, arrival=r.arrival, nsteps=r.nsteps, domain=r.domain, seed=r.seed)
            for r in reqs]


def run_served(rate: float, duration: float, npool: int, budget_gb: float, seed: int = 0,
               serial_seconds: float = None) -> dict:
    """Served workload: Poisson step requests, each carrying its domain's
    state (uploaded inside its latency), against the continuous-batching
    server (FP32-store); and the same requests -- those arriving in the first
    ``serial_seconds`` (default: all) -- served one at a time by the plain
    FP64 single-domain entry, in real time.  With the serial baseline on a
    prefix, the served figures over that same prefix are reported beside it
    (``prefix``)."""
    import torch
    from more_benchmarks_dc_ocean_amd import serve
    from more_benchmarks_dc_ocean_amd.autotune import TUNED, TuneCache

    dev = "cuda:0"
    t0 = time.time()
    cases = serve.pool_cases(npool, seed)
    pool32 = [serve.PinnedDomain(c, "fp32_store") for c in cases]
    log(f"[served] pool of {npool} domains generated and pinned in {time.time() - t0:.1f}s")
    cache = TuneCache(TUNED)
    shapes = [cache.batch_shape(c.nx, c.ny, c.nz, "fp32_store") for c in cases]
    srv = serve.RediServer(pool32, int(budget_gb * 1e9), "fp32_store", max_slots=256, shapes=shapes,
                           device=dev)
    # warm the graph and the kernels on one request of each domain
    srv.run([serve.Request(rid=i, arrival=0.0, nsteps=1, domain=i) for i in range(npool)])
    reqs = serve.poisson_requests(rate, duration, npool, seed)
    log(f"[served] {len(reqs)} requests over {duration:.0f}s at {rate:.0f}/s")
    st = srv.run(reqs, log=log)
    st.update({"rate_per_s": rate, "duration_s": duration, "pool_domains": npool,
               "budget_GB": budget_gb, "iterations": srv.iterations, "max_active": srv.max_active,
               "form": "fp32_store"})
    srv.close()
    del pool32
    torch.cuda.empty_cache()
    log(f"[served] batched p50 {st['p50_ms']:.1f} ms p99 {st['p99_ms']:.1f} ms "
        f"{st['value']:.1f} Mcell-steps/s")
    rnd = lambda d: {k: (round(v, 3) if isinstance(v, float) else v) for k, v in d.items()}
    out = rnd(st)
    pool64 = [serve.PinnedDomain(c, "fp64") for c in cases]
    serve.run_serial([serve.Request(rid=i, arrival=0.0, nsteps=1, domain=i)
                      for i in range(npool)], pool64, dev)  # warm-up
    window = duration if serial_seconds is None else min(serial_seconds, duration)
    sub = [r for r in reqs if r.arrival < window]
    base = serve.run_serial(_copy_requests(sub), pool64, dev, log=log)
    base.update({"form": "fp64", "arrival_window_s": window})
    del pool64
    torch.cuda.empty_cache()
    log(f"[served] serial FP64 ({len(sub)} requests arriving in {window:.0f}s) p50 "
        f"{base['p50_ms']:.1f} ms p99 {base['p99_ms']:.1f} ms {base['value']:.1f} Mcell-steps/s")
    out["serial_fp64"] = rnd(base)
    if len(sub) < len(reqs):
        lat = np.array([r.latency for r in sub])
        out["prefix"] = {"arrival_window_s": window, "requests": len(sub),
                         "p50_ms": round(float(np.percentile(lat, 50) * 1e3), 3),
                         "p99_ms": round(float(np.percentile(lat, 99) * 1e3), 3)}
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="two_fronts_1440x1080x75", choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--form", default="both", choices=["both", "fp64", "fp32_store"])
    ap.add_argument("--batched", type=int, default=96,
                    help="domains of the batched leg (0 = skip it)")
    ap.add_argument("--serve-rate", type=float, default=200.0)
    ap.add_argument("--serve-seconds", type=float, default=60.0,
                    help="arrival window of the served leg (0 = skip it)")
    ap.add_argument("--serve-serial-seconds", type=float, default=20.0,
                    help="arrival window of the requests the serial FP64 baseline serves (0 = all)")
    ap.add_argument("--serve-sustained-rate", type=float, default=50.0,
                    help="request rate of the second, sustainable served point (0 = skip it)")
    ap.add_argument("--serve-sustained-seconds", type=float, default=20.0)
    ap.add_argument("--serve-pool", type=int, default=24)
    ap.add_argument("--serve-budget-gb", type=float, default=200.0)
    ap.add_argument("--reference-steps", type=int, default=20,
                    help="timed steps of the reference's own benchmark configuration (0 = skip it)")
    ap.add_argument("--accuracy-steps", type=int, default=20,
                    help="steps of the FP32-store accuracy runs beside the FP64 form (0 = skip them)")
    ap.add_argument("--gabight-steps", type=int, default=10,
                    help="timed steps of the real-bathymetry workload (0 = skip it)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark runs on exactly one GPU; the Redi step of one "
                                   "domain is not split across GPUs (--gpus must be 1)"}))
        return 2

    from more_benchmarks_dc_ocean_amd import inputs

    builder, wname = WORKLOADS[args.workload]
    t0 = time.time()
    case = getattr(inputs, builder)(nsteps=args.steps)
    log(f"generated {wname} ({case.nx}x{case.ny}x{case.nz}, nt={case.nt}, "
        f"land={case.land_fraction:.3f}) in {time.time() - t0:.1f}s")
    res = {}
    for form in (["fp64", "fp32_store"] if args.form == "both" else [args.form]):
        res[form] = time_form(case, form, args.steps, args.warmup, args.workload)
    main_form = "fp64" if "fp64" in res else args.form
    m = res[main_form]

    cpu = None
    if not args.no_cpu_baseline:
        try:
            cpu = cpu_baseline(case)
            log(f"cpu baseline: {cpu}")
        except Exception as e:
            log(f"cpu baseline failed: {e}")

    line = {"metric": "Redi Mcell-steps/s", "value": m["value"], "unit": "Mcell-steps/s",
            "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": m["ms_per_step"], "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f64" if main_form == "fp64" else "f32-store/f64",
            "data": "synthetic",
            "config": {"workload": wname, "nx_phys": case.nx_phys, "ny_phys": case.ny_phys,
                       "nz": case.nz, "ntracers": case.nt, "eos": "roquet_spv",
                       "land_fraction": round(case.land_fraction, 4), "form": main_form,
                       "parallelism": "single_gpu", "launch_shape": m["launch_shape"]},
            "roofline": m["roofline"], "cpu_baseline": cpu}
    if "fp64" in res and "fp32_store" in res:
        f32 = dict(res["fp32_store"])
        f32["ratio_vs_fp64"] = round(f32["value"] / res["fp64"]["value"], 4)
        if args.accuracy_steps > 0:
            f32["accuracy"] = fp32_accuracy(case, args.accuracy_steps)
        line["forms"] = {"fp32_store": f32}
    if args.reference_steps > 0:
        line["reference_benchmark"] = time_reference_config(args.reference_steps)
    if args.gabight_steps > 0:
        line["real_bathymetry"] = time_gabight(args.gabight_steps, args.warmup)
    if args.batched > 0:
        from more_benchmarks_dc_ocean_amd import inputs as _inp
        t0 = time.time()
        bcases = _inp.case_batch(args.batched, seed=0, nsteps=args.steps)
        log(f"[batched] generated {len(bcases)} domains in {time.time() - t0:.1f}s")
        line["batched"] = time_batched(bcases, "fp32_store", args.steps, args.warmup)
        del bcases
    if args.serve_seconds > 0:
        line["served"] = run_served(args.serve_rate, args.serve_seconds, args.serve_pool,
                                    args.serve_budget_gb,
                                    serial_seconds=args.serve_serial_seconds or None)
    if args.serve_sustained_rate > 0 and args.serve_sustained_seconds > 0:
        line["served_sustained"] = run_served(args.serve_sustained_rate, args.serve_sustained_seconds,
                                              args.serve_pool, args.serve_budget_gb, seed=1)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
