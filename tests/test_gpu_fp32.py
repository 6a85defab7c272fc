"""FP32-store form against the FP64 oracle (MI355X).

The FP32-store form keeps h, the tracers, the face coefficients and the
sublayer fluxes in FP32 and forms every difference of nearly equal values in
FP64.  Both sides here start from the same inputs: h and hTr rounded to FP32
once, then run through the FP32-store kernels and, widened to FP64, through
the oracle.  Errors, per tracer:

  tendency   dTr / areaT of one step; the product writes it in FP64 beside
             the FP32 state (redi_set_tendency_output), the oracle's is
             hTr1 - hTr0.  Field-relative max|d32 - d64| / max|d64| must be
             <= TEND_TOL = 1e-5 (tendencies are ~0 in most cells, so a
             per-cell relative measure is meaningless there).
  hTr1       the FP32 state after one step against the oracle's FP64 state,
             per cell |a-b|/max(|a|,|b|) <= FIELD_TOL = 2e-7 (FP32 storage
             rounding, 6e-8, plus the tendency error).

Every other field's max / RMS error (PoL, PoR, hEff, Ko mismatches) is
printed beside them.
"""
from __future__ import annotations

import dataclasses

import numpy as np
import pytest

from conftest import golden_names, load_golden, rel_err

import redi_oracle
from more_benchmarks_dc_ocean_amd import inputs

pytestmark = pytest.mark.gpu
TEND_TOL = 1e-5
FIELD_TOL = 2e-7


def fp32_inputs(case):
    """The case with h and hTr rounded to FP32 (what both forms then read)."""
    return dataclasses.replace(case, h=case.h.astype(np.float32).astype(np.float64),
                               hTr=case.hTr.astype(np.float32).astype(np.float64))


def run_fp32(case, device, tiles=None, group=None):
    import torch
    from more_benchmarks_dc_ocean_amd.redi import make_redi
    print(f"[gpu] fp32_store {case.name} {case.nx}x{case.ny}x{case.nz} nt={case.nt}", flush=True)
    r, dom = make_redi(case, device, form="fp32_store")
    if tiles:
        r.set_tiles(*tiles)
    if group:
        r.set_tracer_group(group)
    tend = torch.zeros(dom.hTr.shape, dtype=torch.float64, device=device)
    r.set_tendency_output(tend)
    assert r.calc_coeffs(dom) == 0
    assert r.apply_flux(dom) == 0
    torch.cuda.synchronize()
    out = {k: v.cpu().numpy() for k, v in r.coeffs().items()}
    out["tend"] = tend.cpu().numpy()
    out["hTr1"] = dom.tracers_host()
    r.set_tendency_output(None)
    r.close()
    return out


def error_table(got, orc, case):
    """{field: (max err, rms err)}: field-relative for tendencies, absolute
    for positions, relative to max|hEff| for hEff; Ko as mismatch counts."""
    t64 = orc["hTr1"] - case.hTr
    rows = {}
    for t in range(case.nt):
        e = got["tend"][t] - t64[t]
        # relative to the field's largest / RMS tendency; absolute where the
        # oracle's tendency is zero everywhere
        m = np.abs(t64[t]).max() or 1.0
        rm = np.sqrt(np.mean(t64[t] ** 2)) or 1.0
        rows[f"tend[{t}]"] = (float(np.abs(e).max() / m), float(np.sqrt(np.mean(e ** 2)) / rm))
        pc, _ = rel_err(got["hTr1"][t], orc["hTr1"][t])
        rows[f"hTr1[{t}]"] = (pc, float(np.sqrt(np.mean(((got["hTr1"][t] - orc["hTr1"][t]) /
                                                         np.maximum(np.abs(orc["hTr1"][t]), 1e-30)) ** 2))))
    for g in "uv":
        for k in ("PoL", "PoR"):
            d = np.abs(got[g + k] - orc[g + k])
            rows[g + k] = (float(d.max()), float(np.sqrt(np.mean(d ** 2))))
        d = np.abs(got[g + "hEff"] - orc[g + "hEff"])
        m = np.abs(orc[g + "hEff"]).max() or 1.0
        rows[g + "hEff"] = (float(d.max() / m), float(np.sqrt(np.mean(d ** 2)) / m))
        for k in ("KoL", "KoR"):
            rows[g + k] = (int((got[g + k] != orc[g + k]).sum()), got[g + k].size)
    for k, (a, b) in rows.items():
        print(f"[gpu]   {k:9s} max {a:.3e}  rms {b:.3e}" if not k[1:].startswith("Ko")
              else f"[gpu]   {k:9s} mismatches {a} of {b}", flush=True)
    return rows


def check(case, device):
    q = fp32_inputs(case)
    got = run_fp32(q, device)
    orc = redi_oracle.run_oracle(q, nsteps=1)
    rows = error_table(got, orc, q)
    for t in range(q.nt):
        assert rows[f"tend[{t}]"][0] <= TEND_TOL, f"{case.name}: tracer {t} tendency {rows[f'tend[{t}]']}"
        assert rows[f"hTr1[{t}]"][0] <= FIELD_TOL, f"{case.name}: tracer {t} state {rows[f'hTr1[{t}]']}"
    return rows


@pytest.mark.parametrize("name", golden_names())
def test_fp32_store_against_oracle_fixtures(name, gpu_device):
    case, _ = load_golden(name)
    check(case, gpu_device)


@pytest.mark.parametrize("seed,eos,nz", [
    (31, inputs.EOS_ROQUET_SPV, 50),
    (32, inputs.EOS_WRIGHT97, 25),
    (33, inputs.EOS_LINEAR, 75),
])
def test_fp32_store_random_cases(seed, eos, nz, gpu_device):
    c = inputs.case_fractal(40 + seed, 37, 29, nz, 3, 3, seed=seed, land=0.3,
                            fronts=[("y", 0.5, 1.0, 80_000.0)], noise_sigma=2.0, noise_T=0.5,
                            noise_S=0.05, eos_variant=eos, vanished=(seed % 2 == 1))
    check(c, gpu_device)


def test_fp32_store_front_473x297x75(gpu_device):
    """One step of the 473x297x75 Roquet front (4 tracers, 10.5 M cells)."""
    redi_oracle.set_threads(16)
    check(inputs.case_front(nsteps=1), gpu_device)


@pytest.mark.parametrize("tiles,group", [((64, 1), 1), ((32, 2), 2), ((64, 1), 3)])
def test_fp32_store_launch_shapes_same_bits(tiles, group, gpu_device):
    """Launch shape and tracer grouping change the schedule, not the bits."""
    case, _ = load_golden("roquet_two_fronts_mini")
    q = fp32_inputs(case)
    a = run_fp32(q, gpu_device)
    b = run_fp32(q, gpu_device, tiles=tiles, group=group)
    for k in a:
        assert np.array_equal(a[k], b[k]), k
