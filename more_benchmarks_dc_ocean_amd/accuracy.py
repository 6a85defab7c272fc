"""Error of the FP32-store form against the plain FP64 form, field by field.

The FP64 form reproduces the reference's bits (the oracle and the reference
fixtures pin it), so it stands in for the oracle at sizes the oracle cannot
finish quickly.  Three runs of the same workload:

  A  FP64 form on the FP64 inputs           (the reference result)
  B  FP64 form on the inputs rounded to FP32 (h, tracers)
  C  FP32-store form on the same rounded inputs

and, per field, the max and RMS of the difference, absolute and relative to
the reference field's max / RMS:

  tend[t]     tracer t's step-1 tendency dTr / areaT: C vs B (the form's own
              error, same inputs), C vs A (including the input rounding) and
              B vs A (the FP64 form's own response to rounding its inputs:
              the reference's down-gradient guard is a sign test, so a
              rounding-sized change of T can switch a sublayer's flux on or
              off)
  T, S        layer temperature / salinity hTr / h after N steps: C vs A
  hTr[t]      every tracer after N steps: C vs A
  PoL..vhEff  the step-1 Phase-A coefficients: C vs B; KoL/KoR as the
              fraction of (face, surface) entries whose layer differs
  guard       per tracer, the step-1 sublayer fluxes the down-gradient
              guard (redi/ocean_redi.F90:913-924) zeroed in each run, and
              how many of those decisions differ between the runs
              (redi_set_guard_output)

Run: python -m more_benchmarks_dc_ocean_amd.accuracy [--steps 50] [--out file.md]
"""
from __future__ import annotations

import argparse
import dataclasses
import json
import sys
import time

import numpy as np

from . import inputs
from .redi import COEFF_NAMES, make_redi


def _run(case, form: str, nsteps: int, dev: str, want_coeffs: bool):
    import torch
    r, dom = make_redi(case, dev, form)
    tend = torch.zeros(dom.hTr.shape, dtype=torch.float64, device=dev)
    r.set_tendency_output(tend)
    nch = r.guard_chunks()
    gu = torch.zeros((case.nt, (case.nx + 1) * case.ny, nch), dtype=torch.int64, device=dev)
    gv = torch.zeros((case.nt, case.nx * (case.ny + 1), nch), dtype=torch.int64, device=dev)
    r.set_guard_output(gu, gv)
    r.step(dom, 1)
    r.set_tendency_output(None)
    r.set_guard_output(None, None)
    guard = [torch.cat([gu[t].flatten(), gv[t].flatten()]) for t in range(case.nt)]
    co = r.coeffs() if want_coeffs else None
    r.step(dom, nsteps - 1)
    torch.cuda.synchronize()
    h = dom.h.double()
    out = {"tend": tend, "hTr": dom.hTr.double(), "h": h, "coeffs": co, "guard": guard}
    r.close()
    del dom
    torch.cuda.empty_cache()
    return out


def _err(a, ref) -> dict:
    """max / RMS of a - ref, absolute and relative to ref's max / RMS."""
    import torch
    d = (a.double() - ref.double())
    ref = ref.double()
    dmax = float(d.abs().max())
    drms = float(torch.sqrt((d * d).mean()))
    rmax = float(ref.abs().max())
    rrms = float(torch.sqrt((ref * ref).mean()))
    return {"max_abs": dmax, "rms_abs": drms, "max_rel": dmax / rmax if rmax else 0.0,
            "rms_rel": drms / rrms if rrms else 0.0}


def report(case, nsteps: int, dev: str = "cuda:0", log=print) -> dict:
    import torch
    rounded = dataclasses.replace(case, h=case.h.astype(np.float32).astype(np.float64),
                                  hTr=case.hTr.astype(np.float32).astype(np.float64))
    t0 = time.time()
    A = _run(case, "fp64", nsteps, dev, False)
    log(f"A (fp64, fp64 inputs) done {time.time() - t0:.1f}s")
    B = _run(rounded, "fp64", nsteps, dev, True)
    log(f"B (fp64, rounded inputs) done {time.time() - t0:.1f}s")
    C = _run(rounded, "fp32_store", nsteps, dev, True)
    log(f"C (fp32_store) done {time.time() - t0:.1f}s")
    names = ["T", "S"] + [f"tracer{t + 1}" for t in range(2, case.nt)]
    res = {"workload": case.name, "steps": nsteps, "fields": {}}
    F = res["fields"]
    for t, n in enumerate(names):
        F[f"tend_{n} (C vs B)"] = _err(C["tend"][t], B["tend"][t])
        F[f"tend_{n} (C vs A)"] = _err(C["tend"][t], A["tend"][t])
        F[f"tend_{n} (B vs A)"] = _err(B["tend"][t], A["tend"][t])
        F[f"hTr_{n} after {nsteps} (C vs A)"] = _err(C["hTr"][t], A["hTr"][t])
    wet = A["h"] > 0
    for t, n in enumerate(names[:2]):
        la = torch.where(wet, A["hTr"][t] / A["h"].clamp_min(1e-20), torch.zeros_like(A["h"]))
        lc = torch.where(wet, C["hTr"][t] / C["h"].clamp_min(1e-20), torch.zeros_like(C["h"]))
        F[f"{n} layer value after {nsteps} (C vs A)"] = _err(lc[wet], la[wet])
    for k in COEFF_NAMES:
        a, b = C["coeffs"][k], B["coeffs"][k]
        if "Ko" in k:
            F[f"{k} (C vs B)"] = {"mismatch_frac": float((a != b).double().mean())}
        else:
            F[f"{k} (C vs B)"] = _err(a, b)
    res["guard"] = guard_flips(A["guard"], B["guard"], C["guard"], names)
    return res


def _popcount(x) -> int:
    return int(np.bitwise_count(x.cpu().numpy().view(np.uint64)).sum())


def guard_flips(ga, gb, gc, names) -> dict:
    """Per tracer: sublayers the guard zeroed in runs A, B, C and the
    decisions that differ between them (popcounts of XORed bit masks)."""
    out = {}
    for t, n in enumerate(names):
        out[n] = {"zeroed_A": _popcount(ga[t]), "zeroed_B": _popcount(gb[t]),
                  "zeroed_C": _popcount(gc[t]),
                  "flips_B_vs_A": _popcount(ga[t] ^ gb[t]),
                  "flips_C_vs_A": _popcount(ga[t] ^ gc[t]),
                  "flips_C_vs_B": _popcount(gb[t] ^ gc[t])}
    return out


def to_markdown(res: dict, cmd: str) -> str:
    lines = [f"# FP32-store form: error against the plain FP64 form, {res['workload']}", "",
             f"Command: `{cmd}`", "",
             "A = FP64 form on the FP64 inputs; B = FP64 form on the inputs rounded to FP32; "
             "C = FP32-store form on the rounded inputs.  Relative errors are to the reference "
             "field's max (max_rel) and RMS (rms_rel).  C vs B is the form's own error; B vs A "
             "is how much the FP64 step itself moves when its inputs are rounded to FP32 (the "
             "down-gradient guard is a sign test), which C vs A inherits.", "",
             "| field | max abs | RMS abs | max rel | RMS rel |", "|---|---|---|---|---|"]
    for k, e in res["fields"].items():
        if "mismatch_frac" in e:
            lines.append(f"| {k} | layer index differs in {e['mismatch_frac']:.3e} of entries | | | |")
        else:
            lines.append(f"| {k} | {e['max_abs']:.3e} | {e['rms_abs']:.3e} | {e['max_rel']:.3e} | "
                         f"{e['rms_rel']:.3e} |")
    if res.get("guard"):
        lines += ["", "Down-gradient guard decisions of step 1 (sublayer fluxes zeroed, and "
                  "decisions that differ between runs):", "",
                  "| tracer | zeroed A | zeroed B | zeroed C | differ B vs A | differ C vs A | "
                  "differ C vs B |", "|---|---|---|---|---|---|---|"]
        for n, g in res["guard"].items():
            lines.append(f"| {n} | {g['zeroed_A']} | {g['zeroed_B']} | {g['zeroed_C']} | "
                         f"{g['flips_B_vs_A']} | {g['flips_C_vs_A']} | {g['flips_C_vs_B']} |")
    return "\n".join(lines) + "\n"


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--workload", default="two_fronts_1440x1080x75",
                    choices=["two_fronts_1440x1080x75", "front_473x297x75"])
    ap.add_argument("--out", default="")
    args = ap.parse_args()
    builder = {"two_fronts_1440x1080x75": inputs.case_two_fronts,
               "front_473x297x75": inputs.case_front}[args.workload]
    case = builder(nsteps=args.steps)
    log = lambda m: print(f"[accuracy] {m}", file=sys.stderr, flush=True)
    res = report(case, args.steps, log=log)
    md = to_markdown(res, "python -m more_benchmarks_dc_ocean_amd.accuracy " + " ".join(sys.argv[1:]))
    if args.out:
        with open(args.out, "w") as f:
            f.write(md)
    print(md)
    print(json.dumps(res))
    return 0


if __name__ == "__main__":
    sys.exit(main())
