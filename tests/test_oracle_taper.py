"""The opt-in neutral-slope taper of the oracle, pinned by closed forms (CPU).

The reference tapers nothing, so no reference output pins the taper; these
checks do instead.  Two columns of equal, uniform layers with a linear
equation of state, constant salinity and a linear temperature profile, the
right one shifted down by dz: every neutral surface sits dz deeper on the
right, so every interior sublayer has slope dz / dx.  The taper must leave
hEff untouched where dz / dx <= slope_max and scale it by exactly
(slope_max * dx / dz)^2 where dz / dx > slope_max; a mistake in the slope
(a dropped Pa -> m factor, dx in place of 1 / dx, a signed instead of an
absolute slope) or in the taper law changes the ratio.
"""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

import redi_oracle as ro

H = 10.0         # layer thickness, m
NZ = 24
GRAD = 0.02      # dT/dz, K per m
EOS = ro.OracleEos(1, 0, 1035.0, 10.0, 35.0, 2e-4, 7.6e-4, 0.0)  # linear


def _columns(dz):
    """h, T*h, S*h of two columns (top-down layers, fed native bottom-up)."""
    zc = (np.arange(NZ) + 0.5) * H  # layer-centre depth, top-down
    tL = 20.0 - GRAD * zc
    tR = 20.0 - GRAD * (zc - dz)    # the same profile dz deeper
    h = np.full(NZ, H)
    s = np.full(NZ, 35.0)
    flip = lambda a: np.ascontiguousarray(a[::-1])  # native k is bottom-up
    return flip(h), flip(tL * h), flip(s * h), flip(h), flip(tR * h), flip(s * h)


def _coeffs(dz, inv_dx, slope_max):
    ns = 2 * NZ + 2
    PoL, PoR, hEff = np.zeros(ns), np.zeros(ns), np.zeros(ns - 1)
    KoL, KoR = np.zeros(ns, np.int32), np.zeros(ns, np.int32)
    cols = _columns(dz)
    ro.lib().oracle_face_coeffs_taper(NZ, *(ro.ptr(c) for c in cols), ctypes.byref(EOS),
                                      inv_dx, slope_max, ro.ptr(PoL), ro.ptr(PoR), ro.ptr(KoL),
                                      ro.ptr(KoR), ro.ptr(hEff))
    return PoL, PoR, KoL, KoR, hEff


def _interior(KoL, KoR, hEff):
    """Contributing sublayers whose bounding surfaces lie clear of the top and
    bottom layers on both sides (where the PPM edges are not exact)."""
    s = np.arange(hEff.size)
    ok = (hEff != 0)
    for K in (KoL, KoR):
        ok &= (K[s] >= 3) & (K[s + 1] >= 3) & (K[s] <= NZ - 2) & (K[s + 1] <= NZ - 2)
    return ok


@pytest.mark.parametrize("dz,dx", [(3.0, 100.0), (7.5, 250.0), (-4.0, 100.0)])
def test_taper_scales_steep_sublayers_by_the_closed_form(dz, dx):
    slope = abs(dz) / dx
    smax = slope / 3.0
    PoL, PoR, KoL, KoR, plain = _coeffs(dz, 1.0 / dx, 0.0)
    _, _, KoL2, KoR2, tapered = _coeffs(dz, 1.0 / dx, smax)
    assert np.array_equal(KoL, KoL2) and np.array_equal(KoR, KoR2)  # positions untouched
    m = _interior(KoL, KoR, plain)
    assert m.sum() >= NZ  # most sublayers are interior
    ratio = tapered[m] / plain[m]
    assert np.allclose(ratio, (smax / slope) ** 2, rtol=1e-9, atol=0)
    assert np.allclose(ratio, 1.0 / 9.0, rtol=1e-9)


def test_taper_leaves_gentle_slopes_and_off_switch_bitwise():
    dz, dx = 3.0, 100.0
    _, _, KoL, KoR, plain = _coeffs(dz, 1.0 / dx, 0.0)
    gentle = _coeffs(dz, 1.0 / dx, 2.0 * dz / dx)[4]  # interior slopes below the limit
    m = _interior(KoL, KoR, plain)
    assert np.array_equal(plain[m], gentle[m])
    # the sublayers touching the surface / bed are steeper (the columns start
    # and end at the same depths), so only there may the taper act
    assert np.all(gentle[~m] <= plain[~m])
    assert np.array_equal(plain, _coeffs(dz, 1.0 / dx, 1e9)[4])  # no slope above the limit
    ref = np.zeros_like(plain)
    ns = 2 * NZ + 2
    PoL, PoR = np.zeros(ns), np.zeros(ns)
    KoL, KoR = np.zeros(ns, np.int32), np.zeros(ns, np.int32)
    cols = _columns(dz)  # kept alive across the call
    ro.lib().oracle_face_coeffs(NZ, *(ro.ptr(c) for c in cols), ctypes.byref(EOS),
                                ro.ptr(PoL), ro.ptr(PoR), ro.ptr(KoL), ro.ptr(KoR), ro.ptr(ref))
    assert np.array_equal(plain, ref)  # slope_max = 0 is the plain restatement


def test_taper_flat_isopycnals_untouched():
    flat = _coeffs(0.0, 1.0 / 100.0, 1e-6)[4]
    plain = _coeffs(0.0, 1.0 / 100.0, 0.0)[4]
    assert np.array_equal(flat, plain)


def test_taper_reduces_the_step_flux_only_where_steep():
    """Whole step on a fractal domain: with the taper the tendency is never
    larger in magnitude, summed over the domain, than without it."""
    from more_benchmarks_dc_ocean_amd import inputs
    c = inputs.case_domain(3, 55, 20, 18, 25, nt=2)
    a = ro.run_oracle(c, nsteps=1)
    b = ro.run_oracle(c, nsteps=1, slope_max=1e-3)
    da = np.abs(a["hTrN"] - c.hTr).sum()
    db = np.abs(b["hTrN"] - c.hTr).sum()
    assert 0 < db < da
    assert np.all(np.abs(b["uhEff"]) <= np.abs(a["uhEff"]))
    assert np.all(np.abs(b["vhEff"]) <= np.abs(a["vhEff"]))
