"""redi_launch_nt_noalloc (the flat entry with its face state and workspace
in a static device array) against redi_launch_nt, through the C ABI.

Two calls in a row on one device, the second on a smaller domain: every
tracer and every coefficient array must hold the same bits as the
allocating entry's on the same inputs (and the oracle's), so the scratch's
leftovers from the first call cannot leak into the second.  A domain too
large for the scratch is refused on the host before any device work.
"""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

from conftest import load_golden, rel_err

import redi_oracle
from more_benchmarks_dc_ocean_amd import inputs

pytestmark = pytest.mark.gpu


def _coeff_tensors(case, device):
    import torch
    nx, ny, ns = case.nx, case.ny, case.ns
    f64, i32 = torch.float64, torch.int32
    shapes = (("uPoL", (ns, ny, nx + 1), f64), ("uPoR", (ns, ny, nx + 1), f64),
              ("uKoL", (ns, ny, nx + 1), i32), ("uKoR", (ns, ny, nx + 1), i32),
              ("uhEff", (ns - 1, ny, nx + 1), f64), ("vPoL", (ns, ny + 1, nx), f64),
              ("vPoR", (ns, ny + 1, nx), f64), ("vKoL", (ns, ny + 1, nx), i32),
              ("vKoR", (ns, ny + 1, nx), i32), ("vhEff", (ns - 1, ny + 1, nx), f64),
              ("khtr_u", (ny, nx + 1), f64), ("khtr_v", (ny + 1, nx), f64))
    cf = {n: torch.zeros(s, dtype=t, device=device) for n, s, t in shapes}
    for n in ("uKoL", "uKoR", "vKoL", "vKoR"):
        cf[n].fill_(1)  # the reference's init value on faces Phase A never visits
    return cf


def _launch(fn, case, device):
    """One flat step of `case` through `fn`; returns (hTr, coefficients) on the host."""
    import torch
    from more_benchmarks_dc_ocean_amd import redi as R
    dom = R.DeviceDomain(case, device)
    cf = _coeff_tensors(case, device)
    P = lambda t: ctypes.c_void_p(t.data_ptr())
    c = R.Coeffs(*(P(cf[n]) for n in ("uPoL", "uPoR", "uKoL", "uKoR", "uhEff", "vPoL", "vPoR",
                                       "vKoL", "vKoR", "vhEff", "khtr_u", "khtr_v")))
    trs = (ctypes.c_void_p * case.nt)(*(t.data_ptr() for t in dom.hTr))
    torch.cuda.synchronize()
    rc = fn(case.nx, case.ny, case.nz, case.dt, case.nghost, case.nx_phys, case.ny_phys,
            *(int(w) for w in case.walls), ctypes.byref(dom.eos), P(dom.h), case.nt,
            ctypes.cast(trs, ctypes.POINTER(ctypes.c_void_p)), case.idx_T, case.idx_S,
            P(dom.wet_u), P(dom.wet_v), ctypes.byref(c), P(dom.khtr_u_ext), P(dom.khtr_v_ext),
            P(dom.dy_cu), P(dom.dx_cv), P(dom.idxCu), P(dom.idyCv), P(dom.areaT),
            ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream))
    assert rc == 0, R.last_error()
    torch.cuda.synchronize()
    return dom.tracers_host(), {n: t.cpu().numpy() for n, t in cf.items()}


def test_noalloc_matches_allocating_entry_twice(gpu_device):
    from more_benchmarks_dc_ocean_amd import redi as R
    L = R.lib()
    big, _ = load_golden("roquet_mini")          # 20x18x75, Roquet, 4 tracers
    small = inputs.case_seamount(10, 8, 12, nt=3, nsteps=1, seed=9)  # then a smaller one
    for case in (big, small):
        print(f"[gpu] noalloc {case.name} {case.nx}x{case.ny}x{case.nz} nt={case.nt}", flush=True)
        h_ref, c_ref = _launch(L.redi_launch_nt, case, gpu_device)
        h_got, c_got = _launch(L.redi_launch_nt_noalloc, case, gpu_device)
        assert np.array_equal(h_got, h_ref), case.name
        for n in c_ref:
            assert np.array_equal(c_got[n], c_ref[n]), (case.name, n)
        o = redi_oracle.OracleRedi(case)
        h, hTr = np.ascontiguousarray(case.h), case.hTr.copy()
        o.step(h, hTr)
        assert rel_err(h_got, hTr)[0] <= 1e-13, case.name


def test_noalloc_refuses_a_domain_beyond_its_scratch(gpu_device):
    """2000 x 2000 x 75 needs far more than the static scratch: refused on
    the host with REDI_ERR_BUDGET, before any launch (the pointers are never
    dereferenced)."""
    from more_benchmarks_dc_ocean_amd import redi as R
    L = R.lib()
    eos = R.Eos(1, 1, 1035.0, 10.0, 35.0, 2e-4, 7.6e-4, 0.0)
    dummy = (ctypes.c_double * 4)()
    p = ctypes.cast(dummy, ctypes.c_void_p)
    c = R.Coeffs(*([p] * 12))
    trs = (ctypes.c_void_p * 2)(p, p)
    rc = L.redi_launch_nt_noalloc(2000, 2000, 75, 900.0, 3, 1994, 1994, 1, 1, 1, 1,
                                  ctypes.byref(eos), p, 2,
                                  ctypes.cast(trs, ctypes.POINTER(ctypes.c_void_p)), 1, 2, p, p,
                                  ctypes.byref(c), p, p, p, p, p, p, p, None)
    assert rc == -6, rc  # REDI_ERR_BUDGET
    assert "static scratch" in R.last_error()
