"""The library's own kernels and host code, built for the CPU (tests/simt).

tests/simt/build_simt.py compiles more_benchmarks_dc_ocean_amd/csrc with g++
under AddressSanitizer + UndefinedBehaviorSanitizer, and with clang++ under
MemorySanitizer, against a fiber execution model of HIP (tests/simt/simt.hpp):
the threads of a workgroup run one at a time between barriers and
cross-lane operations, in forward, reverse or seeded-random order, LDS
starts every workgroup as undefined bytes, and a barrier or cross-lane
operation reached from divergent code is an error.  The driver
(tests/simt/driver.cpp) runs the C ABI on the reference's input format.

Bars:
  * the reference's fixtures (tests/golden) after one step -- the Phase-A
    coefficients and the tracers -- bit for bit, with no sanitizer report;
  * the same bits whatever the lane and workgroup order (FP64 and
    FP32-store);
  * the one-tracer instantiation of the Phase-B kernels (the ``nt1`` build:
    every tracer in a group of its own) gives the fixtures' bits under
    MemorySanitizer;
  * the batched entry, and several steps, give the fixtures' bits;
  * the self-test kernels show each check firing (tests/simt/selftest.cpp).
"""
from __future__ import annotations

import dataclasses
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from conftest import ROOT, golden_names, load_golden

sys.path.insert(0, os.path.join(ROOT, "tests", "simt"))
import build_simt  # noqa: E402

from more_benchmarks_dc_ocean_amd.inputs import read_output, write_input  # noqa: E402

COEFFS = ("uPoL", "uPoR", "uKoL", "uKoR", "uhEff", "vPoL", "vPoR", "vKoL", "vKoR", "vhEff")


@pytest.fixture(scope="session")
def simt():
    jobs = [("base", "asan"), ("nt1", "asan"), ("nt1", "msan")]
    with ThreadPoolExecutor(len(jobs)) as ex:
        libs = list(ex.map(lambda j: build_simt.build(j[0], 1, False, j[1]), jobs))
    out = {f"{v}_{s}": p for (v, s), p in zip(jobs, libs)}
    for s in ("asan", "msan"):
        out["selftest_" + s] = build_simt.build_selftest(s)
    return out


def _run(exe: str, case, tmp_path, *args) -> dict:
    src, dst = str(tmp_path / "in.bin"), str(tmp_path / "out.bin")
    write_input(src, case)
    r = subprocess.run([exe, src, dst, *args], capture_output=True, text=True, timeout=900)
    assert r.returncode == 0, f"{os.path.basename(exe)} {args}: exit {r.returncode}\n{r.stderr[-4000:]}"
    assert "Sanitizer" not in r.stderr and "runtime error" not in r.stderr, r.stderr[-4000:]
    assert ", 0 reads of returned lanes" in r.stderr, r.stderr[-2000:]
    return read_output(dst)


def _one_step(name):
    case, outs = load_golden(name)
    return dataclasses.replace(case, nsteps=1), outs


def _assert_golden(o, outs, keys=COEFFS + ("hTr1",), what=""):
    for k in keys:
        assert np.array_equal(o[k], outs[k]), f"{what}: {k} differs from the reference fixture"


@pytest.mark.parametrize("name", golden_names())
def test_simt_asan_golden_one_step(simt, tmp_path, name):
    """Random lane and workgroup order, ASan + UBSan: the fixture's bits."""
    case, outs = _one_step(name)
    o = _run(simt["base_asan"], case, tmp_path, "policy=2", "seed=5")
    _assert_golden(o, outs, what=name)


@pytest.mark.parametrize("name", golden_names())
def test_simt_msan_one_tracer_kernels(simt, tmp_path, name):
    """Every tracer in a group of its own, run by the one-tracer kernels,
    under MemorySanitizer: the fixture's bits."""
    case, outs = _one_step(name)
    o = _run(simt["nt1_msan"], case, tmp_path, "group=1", "policy=2", "seed=9")
    _assert_golden(o, outs, what=name)


def test_simt_one_tracer_after_a_pair(simt, tmp_path):
    """Three tracers in groups of two: a pair, then a lone tracer through the
    one-tracer kernels (ASan + UBSan)."""
    case, outs = _one_step("open_scalar_mini")
    assert case.nt == 3
    o = _run(simt["nt1_asan"], case, tmp_path, "group=2", "policy=1")
    _assert_golden(o, outs, what="open_scalar_mini, group 2")


@pytest.mark.parametrize("form", [0, 1])
def test_simt_lane_order_does_not_change_bits(simt, tmp_path, form):
    """Lowest-first, highest-first and two random orders give the same bits,
    in both forms (FP32-store: under MemorySanitizer)."""
    case, outs = _one_step("roquet_two_fronts_mini" if form else "open_scalar_mini")
    exe = simt["nt1_msan"] if form else simt["base_asan"]
    runs = [_run(exe, case, tmp_path, f"form={form}", f"policy={p}", f"seed={s}")
            for p, s in ((0, 1), (1, 1), (2, 3), (2, 4))]
    for o in runs[1:]:
        for k in COEFFS + ("hTr1",):
            assert np.array_equal(o[k], runs[0][k]), f"form {form}: {k} depends on the lane order"
    if form == 0:
        _assert_golden(runs[0], outs, what="open_scalar_mini")
    else:
        assert np.isfinite(runs[0]["hTr1"]).all()
        d = np.abs(runs[0]["hTr1"] - outs["hTr1"]).max() / np.abs(outs["hTr1"]).max()
        assert d < 1e-5, f"FP32-store step-1 tracers {d:.2e} (relative to the max) off the fixture"


@pytest.mark.parametrize("name", ["open_scalar_mini", "roquet_two_fronts_mini"])
def test_simt_batched_entry(simt, tmp_path, name):
    """One domain through redi_batch_* (persistent kernels over a work
    table, a captured graph per step): the fixture's bits."""
    case, outs = _one_step(name)
    o = _run(simt["base_asan"], case, tmp_path, "batch=1", "group=4", "policy=2")
    _assert_golden(o, outs, what=f"{name}, batched")


def test_simt_several_steps(simt, tmp_path):
    case, outs = load_golden("nz1_mini")
    assert case.nsteps > 1
    o = _run(simt["base_asan"], case, tmp_path, "policy=1")
    _assert_golden(o, outs, keys=COEFFS + ("hTr1", "hTrN"), what="nz1_mini")


def _selftest(exe, mode, policy=0):
    return subprocess.run([exe, mode, str(policy)], capture_output=True, text=True, timeout=60)


def test_simt_selftest_lane_order_exposes_races(simt):
    exe = simt["selftest_asan"]
    race = [_selftest(exe, "race", p) for p in (0, 1)]
    assert all(r.returncode == 0 for r in race)
    assert race[0].stdout != race[1].stdout, "an unsynchronised LDS exchange gave order-free results"
    synced = [_selftest(exe, "race_sync", p) for p in (0, 1, 2)]
    want = " ".join(str((l + 1) % 64 + 1) for l in range(64))
    assert all(r.stdout.split() == want.split() for r in synced)


def test_simt_selftest_cross_lane_semantics(simt):
    exe = simt["selftest_asan"]
    v = [float(x) for x in _selftest(exe, "shfl").stdout.split()]
    assert v[:64] == [float(l + 1) for l in range(63)] + [63.0]  # lane 63 keeps its own value
    assert v[64:] == [float(63 - l) for l in range(64)]
    c = _selftest(exe, "compact").stdout.split()
    assert int(c[0], 16) == sum(1 << l for l in range(0, 64, 3))
    assert [int(x) for x in c[1:]] == list(range(0, 64, 3))
    r = _selftest(exe, "returned").stdout.split()
    assert r[0] == "31" and r[1] in ("nan", "-nan") and r[2] == "1"


def test_simt_selftest_reports(simt):
    d = _selftest(simt["selftest_asan"], "divergent")
    assert d.returncode == 3 and "divergent barrier" in d.stderr
    o = _selftest(simt["selftest_asan"], "oob")
    assert o.returncode != 0 and ("out of bounds" in o.stderr or "AddressSanitizer" in o.stderr)
    for mode in ("uninit", "lds_uninit"):
        m = _selftest(simt["selftest_msan"], mode)
        assert m.returncode != 0 and "MemorySanitizer: use-of-uninitialized-value" in m.stderr, mode
        assert "k_" + mode in m.stderr, f"{mode}: not reported inside the kernel"
