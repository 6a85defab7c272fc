"""Batched form: many uneven domains per launch (MI355X, through the C ABI).

Bar: every domain's result in a batch is bit-identical to the same domain
stepped alone -- to the oracle for the plain FP64 form, to the single-domain
FP32-store launches for the FP32-store form -- whatever the other domains,
their order of admission or retirement, the per-domain launch shapes and
the persistence of the kernels.  The domains cover all three layer counts
of the batched workload (25, 50, 75: both accumulator depth classes of the
cell-flux kernel), uneven sizes, and 2, 3 and 4 tracers in a 4-slot batch
(the spare slots' results go to a discard field).
"""
from __future__ import annotations

import numpy as np
import pytest

from conftest import rel_err

import redi_oracle
from more_benchmarks_dc_ocean_amd import inputs

pytestmark = pytest.mark.gpu

# (seed, nx_phys, ny_phys, nz, nt, launch shape)
DOMAINS = [
    (101, 24, 17, 25, 4, {"tile_i": 64, "tile_j": 1, "kchunk": 16}),
    (102, 40, 33, 50, 2, {"tile_i": 16, "tile_j": 4, "kchunk": 8}),
    (103, 13, 29, 75, 3, {"tile_i": 8, "tile_j": 8, "kchunk": 32}),
    (104, 31, 31, 25, 4, {"tile_i": 32, "tile_j": 2, "kchunk": 16}),
]


def _cases():
    return [(inputs.case_domain(3, s, nx, ny, nz, nt=nt, nsteps=2), shape)
            for (s, nx, ny, nz, nt, shape) in DOMAINS]


@pytest.fixture(scope="module")
def cases():
    return _cases()


def _batch(form, cases, device, persist=1, arena=None):
    from more_benchmarks_dc_ocean_amd.batch import RediBatch, domain_bytes
    from more_benchmarks_dc_ocean_amd.redi import DeviceDomain
    need = sum(domain_bytes(c.nx, c.ny, c.nz, form, 4) for c, _ in cases)
    b = RediBatch(form=form, group=4, max_domains=max(16, len(cases)),
                  arena_bytes=arena or need + (1 << 20), persist=persist, device=device)
    doms, slots = [], []
    for c, shape in cases:
        d = DeviceDomain(c, device, form)
        slots.append(b.add(d, shape))
        doms.append(d)
    return b, doms, slots


def test_batch_fp64_matches_oracle_bits(cases, gpu_device):
    import torch
    print(f"[gpu] batch fp64: {len(cases)} domains vs oracle", flush=True)
    b, doms, slots = _batch("fp64", cases, gpu_device)
    b.step(1)
    torch.cuda.synchronize()
    got1 = [(d.tracers_host(), {k: v.cpu().numpy() for k, v in b.coeffs(s).items()})
            for d, s in zip(doms, slots)]
    b.step(1)
    torch.cuda.synchronize()
    for (c, _), d, (h1, co) in zip(cases, doms, got1):
        ref = redi_oracle.run_oracle(c, nsteps=2)
        for k, v in co.items():
            assert np.array_equal(v, ref[k]), f"{c.name}: {k} differs from the oracle"
        assert np.array_equal(h1, ref["hTr1"]), f"{c.name}: hTr after step 1 differs"
        hN = d.tracers_host()
        pc, _ = rel_err(hN, ref["hTrN"])
        assert np.array_equal(hN, ref["hTrN"]), f"{c.name}: hTr after step 2 differs ({pc:.2e})"
    b.close()


def _single_fp32(case, device, nsteps):
    import torch
    from more_benchmarks_dc_ocean_amd.redi import make_redi
    r, dom = make_redi(case, device, "fp32_store")
    r.step(dom, nsteps)
    torch.cuda.synchronize()
    out = dom.tracers_host()
    r.close()
    return out


def test_batch_fp32_same_bits_as_single_domain_and_serial(cases, gpu_device):
    import torch
    print(f"[gpu] batch fp32_store: {len(cases)} domains vs single-domain and serial", flush=True)
    b, doms, _ = _batch("fp32_store", cases, gpu_device)
    b.step(2)
    torch.cuda.synchronize()
    batched = [d.tracers_host() for d in doms]
    b.close()
    b, doms, _ = _batch("fp32_store", cases, gpu_device)
    b.step_serial(2)
    torch.cuda.synchronize()
    serial = [d.tracers_host() for d in doms]
    b.close()
    for (c, _), x, y in zip(cases, batched, serial):
        single = _single_fp32(c, gpu_device, 2)
        assert x.dtype == np.float32
        assert np.array_equal(x, single), f"{c.name}: batched != single-domain"
        assert np.array_equal(y, single), f"{c.name}: serial != single-domain"
        assert not np.array_equal(x, c.hTr.astype(np.float32)), f"{c.name}: nothing changed"


@pytest.mark.parametrize("persist", [0, 2])
def test_batch_persistence_same_bits(cases, gpu_device, persist):
    import torch
    print(f"[gpu] batch fp32_store persist={persist}", flush=True)
    b, doms, _ = _batch("fp32_store", cases, gpu_device, persist=1)
    b.step(2)
    torch.cuda.synchronize()
    ref = [d.tracers_host() for d in doms]
    b.close()
    b, doms, _ = _batch("fp32_store", cases, gpu_device, persist=persist)
    b.step(1)
    b.step(1)
    torch.cuda.synchronize()
    for (c, _), d, r in zip(cases, doms, ref):
        assert np.array_equal(d.tracers_host(), r), f"{c.name}: persist={persist} differs"
    b.close()


def test_batch_result_same_alone_and_in_changing_mix(cases, gpu_device):
    """Domain A stepped 3 times alone equals A stepped 3 times while other
    domains join and leave the batch between its steps; a domain that joins
    late equals that domain stepped alone for its own steps."""
    import torch
    from more_benchmarks_dc_ocean_amd.batch import RediBatch, domain_bytes
    from more_benchmarks_dc_ocean_amd.redi import DeviceDomain
    print("[gpu] batch membership changes", flush=True)
    form = "fp32_store"
    (a, sa), (bb, sb), (cc, sc), (dd, sd) = cases
    b, (da,), _ = _batch(form, [(a, sa)], gpu_device)
    b.step(3)
    torch.cuda.synchronize()
    alone_a = da.tracers_host()
    b.close()
    b, (dc,), _ = _batch(form, [(cc, sc)], gpu_device)
    b.step(2)
    torch.cuda.synchronize()
    alone_c = dc.tracers_host()
    b.close()

    need = sum(domain_bytes(c.nx, c.ny, c.nz, form, 4) for c in (a, bb, cc, dd))
    mix = RediBatch(form=form, group=4, max_domains=4, arena_bytes=need, device=gpu_device)
    xa, xb, xd = (DeviceDomain(a, gpu_device, form), DeviceDomain(bb, gpu_device, form),
                  DeviceDomain(dd, gpu_device, form))
    mix.add(xd, sd)
    slot_b = mix.add(xb, sb)
    mix.add(xa, sa)
    mix.step(1)
    xc = DeviceDomain(cc, gpu_device, form)
    mix.add(xc, sc)          # joins after A's first step
    mix.step(1)
    mix.remove(slot_b)       # leaves after A's second step
    mix.step(1)
    torch.cuda.synchronize()
    assert np.array_equal(xa.tracers_host(), alone_a), "A differs in the mix"
    assert np.array_equal(xc.tracers_host(), alone_c), "late joiner C differs in the mix"
    assert mix.info()["domains"] == 3
    mix.close()


def test_batch_budget_and_slots(cases, gpu_device):
    from more_benchmarks_dc_ocean_amd.batch import BudgetError, RediBatch, domain_bytes
    from more_benchmarks_dc_ocean_amd.redi import DeviceDomain
    print("[gpu] batch budget", flush=True)
    (a, sa), (bb, sb) = cases[0], cases[1]
    na = domain_bytes(a.nx, a.ny, a.nz, "fp32_store", 4)
    nb = domain_bytes(bb.nx, bb.ny, bb.nz, "fp32_store", 4)
    assert na % 256 == 0 and nb % 256 == 0
    b = RediBatch(form="fp32_store", group=4, max_domains=2, arena_bytes=max(na, nb),
                  device=gpu_device)
    da, db = DeviceDomain(a, gpu_device, "fp32_store"), DeviceDomain(bb, gpu_device, "fp32_store")
    s = b.add(da, sa)
    assert b.info()["used_bytes"] == na
    with pytest.raises(BudgetError):
        b.add(db, sb)
    b.remove(s)
    assert b.info() == {"domains": 0, "used_bytes": 0, "largest_free_bytes": max(na, nb)}
    b.add(db, sb)
    b.step(1)
    b.close()


def test_slope_taper_matches_oracle_single_and_batched(cases, gpu_device):
    """The opt-in slope taper (no reference counterpart; pinned on the CPU by
    closed forms in test_oracle_taper.py): the FP64 form with the taper on
    reproduces the oracle's tapered step bit for bit, alone and batched."""
    import torch
    from more_benchmarks_dc_ocean_amd.batch import RediBatch, domain_bytes
    from more_benchmarks_dc_ocean_amd.redi import DeviceDomain, make_redi
    print("[gpu] slope taper fp64 vs oracle", flush=True)
    smax = 2e-3
    sub = cases[:2]
    need = sum(domain_bytes(c.nx, c.ny, c.nz, "fp64", 4) for c, _ in sub)
    b = RediBatch(form="fp64", group=4, max_domains=4, arena_bytes=need, device=gpu_device)
    bd = []
    for c, shape in sub:
        d = DeviceDomain(c, gpu_device, "fp64")
        b.add(d, shape, slope_max=smax)
        bd.append(d)
    b.step(2)
    torch.cuda.synchronize()
    for (c, _), d in zip(sub, bd):
        ref = redi_oracle.run_oracle(c, nsteps=2, slope_max=smax)
        plain = redi_oracle.run_oracle(c, nsteps=1)
        assert not np.array_equal(ref["uhEff"], plain["uhEff"]), "the taper never acted"
        r, dom = make_redi(c, gpu_device, "fp64")
        r.set_slope_taper(smax)
        assert r.calc_coeffs(dom) == 0
        co = {k: v.cpu().numpy() for k, v in r.coeffs().items()}
        assert r.apply_flux(dom) == 0
        r.step(dom, 1)
        torch.cuda.synchronize()
        for k in ("uPoL", "uKoL", "uhEff", "vPoR", "vKoR", "vhEff"):
            assert np.array_equal(co[k], ref[k]), f"{c.name}: tapered {k} differs"
        assert np.array_equal(dom.tracers_host(), ref["hTrN"]), f"{c.name}: single tapered step"
        assert np.array_equal(d.tracers_host(), ref["hTrN"]), f"{c.name}: batched tapered step"
        r.close()
    b.close()


def _interior_totals(case, hTr):
    """sum(areaT * hTr) over the interior of every tracer, in FP64."""
    g, nxp, nyp = case.nghost, case.nx_phys, case.ny_phys
    a = case.areaT[g:g + nyp, g:g + nxp]
    return np.array([(hTr[t][:, g:g + nyp, g:g + nxp].astype(np.float64) * a).sum()
                     for t in range(hTr.shape[0])])


def test_batch_96_domain_workload(gpu_device):
    """The batched workload as specified -- inputs.case_batch(96): 96 domains,
    square sides log-uniform in [48, 512], nz in {25, 50, 75}, each with its
    own seeded bathymetry and T/S, Roquet EOS, 4 tracers -- in the
    FP32-store form with the tuned per-domain launch shapes, 2 steps.  For
    every domain: the batched (one hipGraph launch) bits equal the one
    launch sequence per domain's, values stay finite, each tracer's
    interior total is conserved, and T actually changes."""
    import torch
    from more_benchmarks_dc_ocean_amd.autotune import TUNED, TuneCache
    cases = inputs.case_batch(96, seed=0, nsteps=2)
    sides = [c.nx_phys for c in cases]
    assert len(cases) == 96 and min(sides) >= 48 and max(sides) <= 512
    assert {c.nz for c in cases} == {25, 50, 75}
    cache = TuneCache(TUNED)
    shapes = [cache.batch_shape(c.nx, c.ny, c.nz, "fp32_store") for c in cases]
    print(f"[gpu] batch 96 domains, {sum(c.interior_cells for c in cases) / 1e6:.0f} M cells", flush=True)
    b, doms, _ = _batch("fp32_store", list(zip(cases, shapes)), gpu_device)
    b.step(2)
    torch.cuda.synchronize()
    batched = [d.tracers_host() for d in doms]
    for c, d in zip(cases, doms):
        d.set_tracers(c.hTr)
    b.step_serial(2)
    torch.cuda.synchronize()
    for i, (c, d, x) in enumerate(zip(cases, doms, batched)):
        y = d.tracers_host()
        assert np.array_equal(x, y), f"domain {i} {c.name}: batched != one launch sequence per domain"
        assert np.isfinite(x).all(), f"domain {i}: non-finite values"
        h0 = c.hTr.astype(np.float32)
        t0, t1 = _interior_totals(c, h0), _interior_totals(c, x)
        rel = np.abs(t1 - t0) / np.abs(t0)
        assert (rel < 1e-6).all(), f"domain {i}: interior totals moved by {rel}"
        assert np.count_nonzero(x[0] != h0[0]) > 0.01 * x[0].size, f"domain {i}: T did not change"
    b.close()


# six domains of the 96-domain workload: the smallest and the two largest
# sides, and a mid-size domain of each layer count
CFG3_SUBSET = [54, 32, 58, 94, 16, 48]


def test_batch_fp64_subset_of_96_domain_workload_matches_oracle(gpu_device):
    """Six domains of inputs.case_batch(96) (48^2 to 506^2, nz 25 / 50 / 75)
    batched in the plain FP64 form: Phase-A arrays after step 1 and hTr
    after steps 1 and 2 are the oracle's bits."""
    import torch
    every = inputs.case_batch(96, seed=0, nsteps=2)
    cases = [(every[i], None) for i in CFG3_SUBSET]
    del every
    assert {c.nz for c, _ in cases} == {25, 50, 75}
    assert min(c.nx_phys for c, _ in cases) == 48 and max(c.nx_phys for c, _ in cases) >= 490
    print(f"[gpu] batch fp64: 6 domains of the 96-domain workload vs oracle", flush=True)
    b, doms, slots = _batch("fp64", cases, gpu_device)
    b.step(1)
    torch.cuda.synchronize()
    got1 = [(d.tracers_host(), {k: v.cpu().numpy() for k, v in b.coeffs(s).items()})
            for d, s in zip(doms, slots)]
    b.step(1)
    torch.cuda.synchronize()
    redi_oracle.set_threads(16)
    for (c, _), d, (h1, co) in zip(cases, doms, got1):
        ref = redi_oracle.run_oracle(c, nsteps=2)
        for k, v in co.items():
            assert np.array_equal(v, ref[k]), f"{c.name}: {k} differs from the oracle"
        assert np.array_equal(h1, ref["hTr1"]), f"{c.name}: hTr after step 1 differs"
        assert np.array_equal(d.tracers_host(), ref["hTrN"]), f"{c.name}: hTr after step 2 differs"
    b.close()
