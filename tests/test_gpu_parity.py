"""HIP Redi step against the oracle and the reference fixtures (MI355X).

Every call goes through the C ABI of libredi_hip.so.  Bar for the plain FP64
form: the reference's own outputs reproduced to <= 1e-13 per-cell relative
error (the stated tolerance for the plain form; the reference's cross-checks use
1e-12, redi/drivers/dc_main.F90:237-241).  In practice the kernels follow the
reference's operation order with FP contraction off, so the comparisons below
also record whether the match is exact.
"""
from __future__ import annotations

import hashlib
import sys

import numpy as np
import pytest

from conftest import OUTS, golden_names, load_digests, load_golden, nan_case, rel_err, same_with_nan

import redi_oracle
from more_benchmarks_dc_ocean_amd import inputs

pytestmark = pytest.mark.gpu
TOL = 1e-13


def _run_hip(case, nsteps, device, hoist=False, layer_major=False, tiles=None, group=None):
    import torch
    from more_benchmarks_dc_ocean_amd.redi import make_redi
    print(f"[gpu] {case.name} {case.nx}x{case.ny}x{case.nz} nt={case.nt} steps={nsteps}", flush=True)
    r, dom = make_redi(case, device)
    if layer_major:
        r.set_recon_order(True)
    if tiles:
        r.set_tiles(*tiles)
    if group:
        r.set_tracer_group(group)
    out = {}
    for s in range(nsteps):
        assert r.calc_coeffs(dom) == 0
        assert r.apply_flux(dom, hoist=hoist) == 0
        if s == 0:
            torch.cuda.synchronize()
            out.update({k: v.cpu().numpy() for k, v in r.coeffs().items()})
            out["hTr1"] = dom.tracers_host()
    torch.cuda.synchronize()
    out["hTrN"] = dom.tracers_host()
    r.close()
    return out


def _compare(got, ref, label):
    worst = 0.0
    for k in OUTS:
        if k.endswith(("KoL", "KoR")):
            assert np.array_equal(got[k], ref[k]), f"{label}: {k} differs"
            continue
        pc, fr = rel_err(got[k], ref[k])
        worst = max(worst, pc)
        assert pc <= TOL, f"{label}: {k} per-cell rel err {pc:.3e} (field {fr:.3e})"
    exact = all(np.array_equal(got[k], ref[k]) for k in OUTS)
    print(f"[gpu] {label}: max per-cell rel err {worst:.3e} exact={exact}", flush=True)
    return exact


@pytest.mark.parametrize("name", golden_names())
def test_hip_matches_reference_fixture(name, gpu_device):
    case, ref = load_golden(name)
    got = _run_hip(case, case.nsteps, gpu_device)
    _compare(got, ref, name)


def test_hip_hoist_entry_same_bits(gpu_device):
    case, ref = load_golden("roquet_mini")
    got = _run_hip(case, case.nsteps, gpu_device, hoist=True)
    _compare(got, ref, "roquet_mini/hoist")


def _digest(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def _input_digest(case):
    h = hashlib.sha256()
    for a in ("h", "hTr", "dy_cu", "dx_cv", "idxCu", "idyCv", "areaT", "wet_u", "wet_v",
              "khtr_u_ext", "khtr_v_ext"):
        h.update(np.ascontiguousarray(getattr(case, a), dtype="<f8").tobytes())
    return h.hexdigest()


def _assert_digests(got, d, case, label, names=OUTS):
    """Every output in `names` is byte-identical to the reference build's
    (SHA-256 of the little-endian bytes), on the inputs the reference ran
    (the generator's digest pins them)."""
    from concurrent.futures import ThreadPoolExecutor
    assert _input_digest(case) == d["input_sha256"], f"{label}: generated inputs differ"
    with ThreadPoolExecutor(8) as ex:  # hashlib releases the GIL on large buffers
        dig = dict(zip(names, ex.map(lambda k: _digest(got[k]), names)))
    bad = [k for k in names if dig[k] != d["outputs"][k]["sha256"]]
    print(f"[gpu] {label}: {len(names) - len(bad)}/{len(names)} outputs match the reference's "
          f"digests", flush=True)
    assert not bad, f"{label}: outputs differ from the reference build: {bad}"


def test_hip_full_seamount_64x64x25_against_reference_digests(gpu_device):
    d = load_digests()["seamount_64x64x25_10steps"]
    case = inputs.case_seamount(nsteps=10)
    got = _run_hip(case, 10, gpu_device)
    orc = redi_oracle.run_oracle(case)
    _compare(got, orc, "seamount 64x64x25 x10 vs oracle")
    _assert_digests(got, d, case, "seamount 64x64x25 x10")


def test_hip_full_front_473x297x75_one_step(gpu_device):
    """473x297x75, Roquet, 4 tracers: one full step at the workload's size
    against the oracle run live and the reference's digests."""
    d = load_digests()["front_473x297x75_1step"]
    case = inputs.case_front(nsteps=1)
    got = _run_hip(case, 1, gpu_device)
    redi_oracle.set_threads(16)
    orc = redi_oracle.run_oracle(case, nsteps=1)
    _compare(got, orc, "front 473x297x75 x1 vs oracle")
    _assert_digests(got, d, case, "front 473x297x75 x1")


def test_hip_front_473x297x75_100_steps_reference_digests(gpu_device):
    """The 473x297x75 front workload as specified (Roquet, 4 tracers, 100
    steps): the Phase-A arrays and hTr after step 1, and hTr after step 100,
    byte-identical to the reference build's over the same 100 steps."""
    d = load_digests()["front_473x297x75_100steps"]
    case = inputs.case_front(nsteps=100)
    got = _run_hip(case, 100, gpu_device)
    _assert_digests(got, d, case, "front 473x297x75 x100")


def test_hip_two_fronts_1440x1080x75_one_step_reference_digests(gpu_device):
    """The headline 1440x1080x75 two-front workload (Roquet, 4 tracers): the
    Phase-A arrays and hTr after one step byte-identical to the reference
    build's."""
    d = load_digests()["two_fronts_1440x1080x75_1step"]
    case = inputs.case_two_fronts(nsteps=1)
    got = _run_hip(case, 1, gpu_device)
    _assert_digests(got, d, case, "two fronts 1440x1080x75 x1")


@pytest.mark.parametrize("seed,eos,nz,walls,use_ext", [
    (21, inputs.EOS_ROQUET_SPV, 50, (1, 1, 1, 1), 1),
    (22, inputs.EOS_WRIGHT97, 25, (0, 1, 1, 0), 1),
    (23, inputs.EOS_LINEAR, 2, (1, 1, 1, 1), 0),
    (24, inputs.EOS_ROQUET_SPV, 75, (1, 0, 0, 1), 0),
])
def test_hip_matches_oracle_random_cases(seed, eos, nz, walls, use_ext, gpu_device):
    import dataclasses
    c = inputs.case_fractal(30 + seed, 33, 27, nz, 3, 3, seed=seed, land=0.3,
                            fronts=[("y", 0.5, 1.0, 80_000.0)], noise_sigma=2.0, noise_T=0.5,
                            noise_S=0.05, eos_variant=eos, vanished=(seed % 2 == 0))
    c = dataclasses.replace(c, walls=walls, use_ext=use_ext, khtr=750.0)
    got = _run_hip(c, 3, gpu_device)
    orc = redi_oracle.run_oracle(c)
    _compare(got, orc, f"random seed={seed}")


@pytest.mark.parametrize("name", ["roquet_mini", "wright_vanished_mini", "seamount_mini"])
def test_hip_layer_major_reconstruction_same_bits(name, gpu_device):
    """The layer-major storage order of the reconstruction gives the same
    result as the per-column default."""
    case, ref = load_golden(name)
    got = _run_hip(case, case.nsteps, gpu_device, layer_major=True)
    _compare(got, ref, name + "/layer-major")


@pytest.mark.parametrize("tiles,group", [((64, 1), 1), ((8, 8), 2), ((32, 4), 3), ((16, 16), 4),
                                         ((256, 1), 4)])
def test_hip_launch_shapes_same_bits(tiles, group, gpu_device):
    """Tile shape and tracer grouping change only the schedule: every
    combination reproduces the reference bits."""
    case, ref = load_golden("roquet_two_fronts_mini")
    got = _run_hip(case, case.nsteps, gpu_device, tiles=tiles, group=group)
    _compare(got, ref, f"roquet_two_fronts_mini tiles={tiles} group={group}")


@pytest.mark.parametrize("nt", [1, 2, 3])
@pytest.mark.parametrize("group", [1, 2, 3, 4])
def test_hip_tracer_counts_and_groups(nt, group, gpu_device):
    """1..3 tracers under every tracer grouping, including groups of one
    tracer (run through the two-tracer kernels with the second result
    discarded) and the Phase-B scratch sized for fewer than four tracers."""
    import dataclasses
    base, _ = load_golden("roquet_two_fronts_mini")
    case = dataclasses.replace(base, hTr=base.hTr[:nt].copy(), idx_T=1, idx_S=min(2, nt),
                               name=f"roquet_two_fronts_mini_nt{nt}")
    o = redi_oracle.run_oracle(case, nsteps=1)
    for layer_major in (False, True):
        got = _run_hip(case, 1, gpu_device, group=group, layer_major=layer_major)
        assert np.array_equal(got["hTr1"], o["hTr1"]), (nt, group, layer_major)


def test_nonmonotone_faces_reported(gpu_device):
    from more_benchmarks_dc_ocean_amd.redi import make_redi
    case = inputs.case_front(nsteps=1)
    r, dom = make_redi(case, gpu_device)
    r.calc_coeffs(dom)
    nu, nv = r.nonmonotone_faces()
    print(f"[gpu] front 473x297x75: non-monotone faces u={nu} v={nv} of "
          f"{(case.nx + 1) * case.ny + case.nx * (case.ny + 1)}", flush=True)
    assert nu >= 0 and nv >= 0
    r.close()


def test_hip_deterministic(gpu_device):
    case, _ = load_golden("roquet_two_fronts_mini")
    a = _run_hip(case, 2, gpu_device)
    b = _run_hip(case, 2, gpu_device)
    for k in OUTS:
        assert np.array_equal(a[k], b[k]), k


@pytest.mark.parametrize("name", ["roquet_two_fronts_mini", "wright_vanished_mini", "nz1_mini"])
def test_hip_flux_segment_same_bits(name, gpu_device):
    """Runs of 1, 3, 7 and 1000 cells in the flux kernels (every run's end
    faces computed twice, or whole rows and columns in one run): the
    oracle's bits in the FP64 form, the default run's bits in the
    FP32-store form."""
    import torch
    from more_benchmarks_dc_ocean_amd.redi import make_redi
    case, ref = load_golden(name)
    print(f"[gpu] flux segment {name}", flush=True)
    for form in ("fp64", "fp32_store"):
        outs = []
        for seg in (32, 1, 3, 7, 1000):
            r, dom = make_redi(case, gpu_device, form)
            r.set_flux_segment(seg)
            r.step(dom, case.nsteps)
            torch.cuda.synchronize()
            outs.append(dom.tracers_host())
            r.close()
        for o in outs[1:]:
            assert np.array_equal(o, outs[0]), form
        if form == "fp64":
            assert np.array_equal(outs[0], ref["hTrN"])


def test_hip_flux_segment_front_473x297x75(gpu_device):
    import torch
    from more_benchmarks_dc_ocean_amd.redi import make_redi
    case = inputs.case_front(nsteps=1)
    print("[gpu] flux segment 473x297x75", flush=True)
    got = []
    for seg in (32, 5, 64):
        r, dom = make_redi(case, gpu_device, "fp64")
        r.set_flux_segment(seg)
        r.step(dom, 1)
        torch.cuda.synchronize()
        got.append(dom.tracers_host())
        r.close()
    assert np.array_equal(got[0], got[1])
    assert np.array_equal(got[0], got[2])


@pytest.mark.parametrize("form", ["fp64", "fp32_store"])
def test_hip_tile_order_same_bits(form, gpu_device):
    """Vertical numbering of the sweep tiles: same bits (the oracle's in the
    FP64 form)."""
    import torch
    from more_benchmarks_dc_ocean_amd.redi import make_redi
    case, ref = load_golden("roquet_two_fronts_mini")
    print(f"[gpu] tile order {form}", flush=True)
    got = []
    for vertical in (False, True):
        r, dom = make_redi(case, gpu_device, form)
        r.set_tile_order(vertical)
        r.set_tiles(16, 4)
        r.step(dom, case.nsteps)
        torch.cuda.synchronize()
        got.append(dom.tracers_host())
        r.close()
    assert np.array_equal(got[0], got[1])
    if form == "fp64":
        assert np.array_equal(got[1], ref["hTrN"])


@pytest.mark.parametrize("nz", [100, 256])
def test_hip_deep_columns_match_oracle_alone_and_batched(nz, gpu_device):
    """The two deepest layer-capacity classes of the flux kernels (nz up to
    128 and up to 256, the library's limit): a 12x10 Roquet domain with
    vanished layers, 2 steps, alone and in a batch, against the oracle."""
    import torch
    from more_benchmarks_dc_ocean_amd.batch import RediBatch, domain_bytes
    from more_benchmarks_dc_ocean_amd.redi import DeviceDomain
    c = inputs.case_fractal(9, 12, 10, nz, 4, 2, seed=31 + nz, land=0.25,
                            fronts=[("y", 0.5, 1.0, 40_000.0)], noise_sigma=2.0, noise_T=0.5,
                            noise_S=0.05, eos_variant=inputs.EOS_ROQUET_SPV, vanished=True)
    ref = redi_oracle.run_oracle(c, nsteps=2)
    got = _run_hip(c, 2, gpu_device)
    _compare(got, ref, f"deep nz={nz} alone")
    assert all(np.array_equal(got[k], ref[k]) for k in OUTS)
    b = RediBatch(form="fp64", group=4, max_domains=2,
                  arena_bytes=domain_bytes(c.nx, c.ny, c.nz, "fp64", 4), device=gpu_device)
    d = DeviceDomain(c, gpu_device, "fp64")
    b.add(d)
    b.step(2)
    torch.cuda.synchronize()
    assert np.array_equal(d.tracers_host(), ref["hTrN"]), f"deep nz={nz} batched"
    b.close()


def test_hip_nan_input_same_as_oracle(gpu_device):
    """One NaN in the temperature input (the case the reference driver's
    sanity stop catches, redi/drivers/dc_main.F90:149-153).  NaN spreads
    through the stencil and, where a branch compares a NaN's sign (the PLM
    and guard signums use copysign), the NaN's sign bit decides it; x86 and
    the GPU give arithmetic on a NaN different sign bits, so the cells near
    the NaN may take other branches than the oracle's.  Bar: the sanity rule
    flags the result as garbage, as it does the reference's, and every
    array agrees with the oracle bit for bit beyond 2 columns per step from
    the NaN's column."""
    from conftest import away_from
    from more_benchmarks_dc_ocean_amd.redi import sanity
    c, where = nan_case()
    got = _run_hip(c, 2, gpu_device)
    ref = redi_oracle.run_oracle(c)
    assert sanity(got["hTrN"][0]) == "garbage"
    for k in OUTS:
        face = "u" if k.startswith("u") else ("v" if k.startswith("v") else None)
        r = 2 if k in ("hTr1",) or face else 4
        m = away_from(got[k].shape, where, r, face)
        assert same_with_nan(got[k][m], ref[k][m]), k
        assert not np.isnan(got[k][m]).any(), k


def test_hip_wet_masks_changing_between_steps(gpu_device):
    """A dry face's records hold the inert values (Po = hEff = 0, Ko = 1) and
    the sweep leaves them alone while the face stays dry.  Faces that turn dry
    between two Phase-A calls get the inert values again, and faces that turn
    wet get computed ones: the coefficients of each call equal the oracle's
    for that call's masks, bit for bit."""
    import torch
    from more_benchmarks_dc_ocean_amd.redi import make_redi
    base = inputs.case_fractal(41, 36, 28, 25, 2, 1, seed=41, land=0.3,
                               fronts=[("x", 0.5, 1.0, 80_000.0)], noise_sigma=2.0, noise_T=0.5,
                               noise_S=0.05)
    drier = base.copy()
    drier.wet_u[:, ::3] = 0.0  # every third column of u faces turns dry
    drier.wet_v[::4, :] = 0.0  # every fourth row of v faces turns dry
    assert (drier.wet_u < base.wet_u).any() and (drier.wet_v < base.wet_v).any()
    r, dom = make_redi(base, gpu_device)
    print(f"[gpu] wet masks {base.nx}x{base.ny}x{base.nz}: base, drier, base", flush=True)
    h = np.ascontiguousarray(base.h)
    hTr = np.ascontiguousarray(base.hTr)
    for label, c in (("base", base), ("drier", drier), ("base again", base)):
        dom.wet_u.copy_(torch.from_numpy(np.ascontiguousarray(c.wet_u)))
        dom.wet_v.copy_(torch.from_numpy(np.ascontiguousarray(c.wet_v)))
        assert r.calc_coeffs(dom) == 0
        got = {k: v.cpu().numpy() for k, v in r.coeffs().items()}
        o = redi_oracle.OracleRedi(c)
        o.calc_coeffs(h, hTr)
        for k, ref in o.coeffs().items():
            assert np.array_equal(got[k], ref), f"{label}: {k} differs from the oracle"
    r.close()
