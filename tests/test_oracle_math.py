"""Closed-form and invariant checks of the oracle (CPU).

These pin the oracle's pieces by what the mathematics fixes, independently
of the compiled reference: PPM/PLM on linear and degenerate profiles, the
position-interpolation table, the EOS closed forms and finite-difference
derivatives, sweep identities on identical and stably stratified columns,
and the conservation / symmetry / scaling properties of a full step.
"""
from __future__ import annotations

import ctypes
import dataclasses

import numpy as np
import pytest

import redi_oracle as ro
from more_benchmarks_dc_ocean_amd import inputs

L = ro.lib()
P = ro.ptr


# --------------------------------------------------------------------------
# PPM / PLM
# --------------------------------------------------------------------------

def interface_scalar(h, tr):
    h = np.ascontiguousarray(h, dtype=np.float64)
    tr = np.ascontiguousarray(tr, dtype=np.float64)
    e = np.zeros(len(h) + 1)
    L.oracle_interface_scalar(len(h), P(h), P(tr), P(e), 1e-30)
    return e


def test_ppm_edges_of_linear_profile_on_uniform_grid():
    # uniform h, linear T (step D per layer): the PLM slope is D inside and
    # 0 in the two PCM end layers.  On a uniform grid the PPM edge reduces to
    # mean - (p_{k+1} - p_k)/6, so interior edges are the neighbour means,
    # the first/last interior edges are mean -/+ D/6, and the two boundary
    # edges are the end layers' own values.
    nk = 9
    D = 0.5
    h = np.full(nk, 7.0)
    t = 3.0 + D * np.arange(nk)
    e = interface_scalar(h, t)
    mean = 0.5 * (t[:-1] + t[1:])
    assert e[0] == t[0] and e[-1] == t[-1]
    np.testing.assert_allclose(e[2:-2], mean[1:-1], rtol=0, atol=1e-13)
    assert e[1] == pytest.approx(mean[0] - D / 6, abs=1e-13)
    assert e[-2] == pytest.approx(mean[-1] + D / 6, abs=1e-13)


def test_plm_slope_zero_at_extremum_and_next_to_vanished_layer():
    h = np.array([10.0, 10.0, 10.0, 0.0, 10.0, 10.0])
    s = np.array([1.0, 3.0, 2.0, 5.0, 6.0, 7.0])
    d = np.zeros(6)
    L.oracle_plm_diff(6, P(h), P(s), P(d))
    assert d[0] == 0.0 and d[-1] == 0.0  # PCM ends
    assert d[1] == 0.0  # local maximum at layer 1
    assert d[3] == 0.0  # vanished layer: (h4+h3)*(h2+h3) > 0 but d_l, d_r differ in sign
    assert d[4] != 0.0  # monotone interior with thick neighbours


def test_fv_diff_degenerate_thicknesses():
    assert L.oracle_fv_diff(0.0, 0.0, 0.0, 1.0, 2.0, 3.0) == 0.0
    # uniform thickness: centred difference (s_{k+1} - s_{k-1}) / 2 * ... = 1 per layer here
    assert L.oracle_fv_diff(1.0, 1.0, 1.0, 1.0, 2.0, 3.0) == pytest.approx(1.0, abs=1e-15)


def test_ppm_edge_zero_thickness_pair_is_mean():
    assert L.oracle_ppm_edge(1.0, 0.0, 0.0, 1.0, 2.0, 4.0, 0.0, 0.0, 1e-30) == 3.0


# --------------------------------------------------------------------------
# interpolate_position and ppm_ave
# --------------------------------------------------------------------------

def test_interpolate_position_table():
    f = L.oracle_interpolate_position
    assert f(-1.0, 0.0, 3.0, 1.0) == 0.25
    assert f(-2.0, 0.0, -1.0, 1.0) == 1.0  # clamped above
    assert f(1.0, 0.0, 2.0, 1.0) == 0.0  # clamped below
    assert f(1.0, 0.0, 1.0, 1.0) == 0.0  # dRhoPos == dRhoNeg > 0
    assert f(-1.0, 0.0, -1.0, 1.0) == 1.0  # == < 0
    assert f(0.0, 0.0, 0.0, 1.0) == 0.5  # == 0
    assert f(-1.0, 1.0, 1.0, 1.0) == 0.5  # vanished layer
    assert f(-1.0, 2.0, 1.0, 1.0) == 0.5  # inverted positions
    assert f(1.0, 0.0, -1.0, 1.0) == 0.5  # dRhoNeg > dRhoPos


def test_ppm_ave_integrates_the_parabola():
    # mean of the PPM parabola a(x) = aL + x((aR-aL) + a6(1-x)) over [xL, xR]
    aL, aR, am = 1.0, 4.0, 2.0
    a6 = 3.0 * (2.0 * am - (aL + aR))
    par = lambda x: aL + x * ((aR - aL) + a6 * (1.0 - x))
    for xL, xR in ((0.0, 1.0), (0.2, 0.7), (0.1, 0.1 + 1e-3)):
        xs = np.linspace(xL, xR, 20001)
        exact = np.trapezoid(par(xs), xs) / (xR - xL)
        assert L.oracle_ppm_ave(xL, xR, aL, aR, am) == pytest.approx(exact, rel=1e-8)
    assert L.oracle_ppm_ave(0.0, 1.0, aL, aR, am) == pytest.approx(am, rel=1e-15)
    # zero-width: point value of the parabola at xR
    assert L.oracle_ppm_ave(0.3, 0.3, aL, aR, am) == pytest.approx(par(0.3), rel=1e-15)


# --------------------------------------------------------------------------
# equation of state
# --------------------------------------------------------------------------

def eos(variant, **kw):
    d = dict(rho0=1035.0, T_ref=10.0, S_ref=35.0, alpha_T=2e-4, beta_S=7.6e-4, p_ref=0.0)
    d.update(kw)
    return ro.OracleEos(variant, 0, d["rho0"], d["T_ref"], d["S_ref"], d["alpha_T"], d["beta_S"],
                        d["p_ref"])


def derivs(e, T, S, p):
    a, b = ctypes.c_double(), ctypes.c_double()
    L.oracle_eos_specvol_derivs(ctypes.byref(e), T, S, p, ctypes.byref(a), ctypes.byref(b))
    return a.value, b.value


def test_linear_eos_closed_form():
    e = eos(1)
    assert L.oracle_eos_density_point(ctypes.byref(e), 12.0, 36.0, 1e7) == \
        (1035.0 + 7.6e-4 * (36.0 - 35.0)) - 2e-4 * (12.0 - 10.0)
    dt, ds = derivs(e, 12.0, 36.0, 1e7)
    assert dt == 2e-4 / (1035.0 * 1035.0) and ds == -(7.6e-4 / (1035.0 * 1035.0))


@pytest.mark.parametrize("variant", [2, 4])
def test_nonlinear_eos_derivatives_match_finite_differences(variant):
    """dSV/dT and dSV/dS against central differences of 1/rho (the check the
    reference names as the lock on its chain rule, ocean_eos.F90:305-306)."""
    e = eos(variant)
    sv = lambda T, S, p: 1.0 / L.oracle_eos_density_point(ctypes.byref(e), T, S, p)
    for T, S, p in ((2.0, 34.5, 0.0), (15.0, 35.5, 1.0e7), (25.0, 36.2, 3.0e7), (-1.0, 33.0, 5.0e7)):
        dt, ds = derivs(e, T, S, p)
        ht, hs = 1e-4, 1e-4
        fdt = (sv(T + ht, S, p) - sv(T - ht, S, p)) / (2 * ht)
        fds = (sv(T, S + hs, p) - sv(T, S - hs, p)) / (2 * hs)
        assert dt == pytest.approx(fdt, rel=2e-6)
        assert ds == pytest.approx(fds, rel=2e-6)


def test_roquet_density_is_seawater_like():
    e = eos(4)
    rho = L.oracle_eos_density_point(ctypes.byref(e), 10.0, 35.0, 0.0)
    assert 1026.0 < rho < 1028.0
    rho_deep = L.oracle_eos_density_point(ctypes.byref(e), 2.0, 34.7, 4.0e7)
    assert 1045.0 < rho_deep < 1048.0
    dt, ds = derivs(e, 10.0, 35.0, 0.0)
    assert dt > 0 and ds < 0  # warm expands, salt contracts


# --------------------------------------------------------------------------
# neutral sweep
# --------------------------------------------------------------------------

def sweep(nk, Pl, Tl, Sl, dTl, dSl, Pr, Tr, Sr, dTr, dSr):
    ns = 2 * nk + 2
    arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (Pl, Tl, Sl, dTl, dSl, Pr, Tr, Sr, dTr, dSr)]
    PoL, PoR, hEff = np.zeros(ns), np.zeros(ns), np.zeros(ns - 1)
    KoL, KoR = np.zeros(ns, np.int32), np.zeros(ns, np.int32)
    L.oracle_neutral_positions_continuous(nk, *(P(a) for a in arrs), P(PoL), P(PoR), P(KoL),
                                          P(KoR), P(hEff))
    return PoL, PoR, KoL, KoR, hEff


def column(nk, T0, dTdz, h=100.0):
    P_ = np.arange(nk + 1) * (9.80665 * 1035.0 * h)
    T = T0 - dTdz * np.arange(nk + 1)
    S = np.full(nk + 1, 35.0)
    dRdT = np.full(nk + 1, -0.2)
    dRdS = np.full(nk + 1, 0.76)
    return P_, T, S, dRdT, dRdS


@pytest.mark.parametrize("nk", [1, 2, 3, 7])
def test_sweep_identical_columns(nk):
    c = column(nk, 20.0, 1.5)
    PoL, PoR, KoL, KoR, hEff = sweep(nk, *c, *c)
    assert np.array_equal(PoL, PoR) and np.array_equal(KoL, KoR)
    # every interface is matched with itself: the sublayers alternate between
    # zero thickness and a full layer, and together cover the column
    assert hEff.sum() == pytest.approx(c[0][-1] - c[0][0], rel=1e-14)


@pytest.mark.parametrize("shift", [0.3, 1.0, 2.5])
def test_sweep_offset_columns_invariants(shift):
    nk = 6
    l = column(nk, 20.0, 1.0)
    r = column(nk, 20.0 - shift, 1.0)
    PoL, PoR, KoL, KoR, hEff = sweep(nk, *l, *r)
    posL = KoL + PoL
    posR = KoR + PoR
    assert np.all(np.diff(posL) >= 0) and np.all(np.diff(posR) >= 0)  # monotone positions
    assert np.all((PoL >= 0) & (PoL <= 1) & (PoR >= 0) & (PoR <= 1))
    assert np.all((KoL >= 1) & (KoL <= nk) & (KoR >= 1) & (KoR <= nk))
    assert np.all(hEff >= 0)
    # harmonic mean <= arithmetic mean: sum of hEff bounded by the column
    assert hEff.sum() <= l[0][-1] * (1 + 1e-14)


def test_sweep_linear_stratification_finds_exact_neutral_positions():
    """Two linearly stratified columns offset by half a layer in T: each left
    interface's neutral depth on the right lies half a layer up."""
    nk = 5
    l = column(nk, 20.0, 1.0)
    r = column(nk, 19.5, 1.0)
    PoL, PoR, KoL, KoR, hEff = sweep(nk, *l, *r)
    # right column interfaces are lighter by 0.5 layer: where the left
    # column is searched, the crossing sits at fractional position 0.5
    inner = (PoL > 0) & (PoL < 1)
    assert inner.any() and np.allclose(PoL[inner], 0.5, atol=1e-12)


# --------------------------------------------------------------------------
# whole-step properties
# --------------------------------------------------------------------------

def _run(case, nsteps=1):
    return ro.run_oracle(case, nsteps=nsteps)["hTrN"]


@pytest.fixture(scope="module")
def small_case():
    return inputs.case_fractal(40, 26, 22, 20, 3, 1, seed=4, land=0.3,
                               fronts=[("y", 0.5, 1.0, 60_000.0)], noise_sigma=2.0,
                               noise_T=0.5, noise_S=0.05)


def test_interior_total_conserved_with_walls(small_case):
    c = small_case
    g = c.nghost
    out = _run(c, 3)
    for t in range(c.nt):
        a0 = (c.hTr[t] * c.areaT)[:, g:-g, g:-g].sum()
        a1 = (out[t] * c.areaT)[:, g:-g, g:-g].sum()
        assert abs(a1 - a0) <= 1e-13 * abs(a0)
        assert not np.array_equal(out[t], c.hTr[t])


def test_uniform_tracer_is_bit_constant(small_case):
    c = small_case.copy()
    c.hTr[2] = 0.25 * c.h
    out = _run(c, 2)
    assert np.array_equal(out[2], c.hTr[2])


def test_horizontally_uniform_state_has_no_flux():
    c = inputs.case_seamount(12, 10, 8, nt=2, seed=1)
    c = c.copy()
    prof_h = c.h[:, c.ny // 2, c.nx // 2].copy()
    ocean = c.h.sum(0) > 0
    c.h = np.where(ocean[None], prof_h[:, None, None], 0.0) * np.ones_like(c.h)
    for t in range(c.nt):
        prof = c.hTr[t][:, c.ny // 2, c.nx // 2].copy()
        c.hTr[t] = np.where(ocean[None], prof[:, None, None], 0.0) * np.ones_like(c.h)
    out = _run(c, 1)
    assert np.array_equal(out, c.hTr)


def test_doubling_kappa_doubles_the_tendency_bitwise(small_case):
    c = small_case
    c2 = dataclasses.replace(c, khtr_u_ext=2.0 * c.khtr_u_ext, khtr_v_ext=2.0 * c.khtr_v_ext)
    d1 = _run(c) - c.hTr
    d2 = _run(c2) - c.hTr
    # coef scales by an exact power of two, so flux bits double; the
    # tendency only picks up the final hTr rounding
    np.testing.assert_allclose(d2, 2.0 * d1, rtol=0, atol=4e-16 * np.abs(c.hTr).max())


def test_zero_scalar_kappa_is_a_noop(small_case):
    c = dataclasses.replace(small_case, use_ext=0, khtr=0.0)
    out = _run(c)
    assert np.array_equal(out, c.hTr)


def test_x_mirror_symmetry():
    c = inputs.case_seamount(14, 10, 6, nt=2, seed=3)
    m = c.copy()
    m.h = c.h[:, :, ::-1].copy()
    m.hTr = c.hTr[:, :, :, ::-1].copy()
    m.wet_u = c.wet_u[:, ::-1].copy()
    m.wet_v = c.wet_v[:, ::-1].copy()
    a = _run(c)
    b = _run(m)
    np.testing.assert_allclose(b[:, :, :, ::-1], a, rtol=1e-13, atol=1e-13 * np.abs(a).max())


def test_wall_faces_carry_no_flux():
    """With walls, the ghost ring and the interior never exchange tracer."""
    c = inputs.case_seamount(10, 8, 6, nt=2, seed=5)
    g = c.nghost
    c = c.copy()
    c.hTr[0][:, :g, :] *= 3.0  # make the ghost rows very different
    out = _run(c)
    a0 = (c.hTr[0] * c.areaT)[:, g:-g, g:-g].sum()
    a1 = (out[0] * c.areaT)[:, g:-g, g:-g].sum()
    assert abs(a1 - a0) <= 1e-13 * abs(a0)
