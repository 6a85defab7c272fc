// redi_eos.hpp -- equation-of-state point routines for the Redi kernels.
//
// Restates redi/ocean_eos.F90:282-546 (roquet_spv_point, eos_specvol_derivs,
// eos_density_point) for gfx950.  The Roquet et al. (2015) SpV coefficients
// are the published ones with the reference's unit scalings folded in
// (redi/ocean_eos.F90:116-238); integer powers are formed by right-to-left
// binary exponentiation (x^5 = x*x^4, x^6 = x^2*x^4), which is how the
// reference compiler folds `x**n`, so every folded coefficient has the same
// bits as the reference build.
//
// Every routine is templated on the arithmetic type R: the plain FP64 form
// instantiates R = double and must be compiled without FP contraction so it
// rounds like the reference; the precision form instantiates R = float.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

namespace redi {

enum EosVariant : int { kEosLinear = 1, kEosWright97 = 2, kEosRoquetSpV = 4 };

// ocean_eos_t (redi/ocean_eos.F90:239-278), by value.
struct EosParams {
  int variant;
  int ts_convention;
  double rho0, T_ref, S_ref, alpha_T, beta_S, p_ref;
};

namespace eosc {
constexpr double pw2(double x) { return x * x; }
constexpr double pw3(double x) { return pw2(x) * x; }
constexpr double pw4(double x) { return pw2(x) * pw2(x); }
constexpr double pw5(double x) { return x * pw4(x); }
constexpr double pw6(double x) { return pw2(x) * pw4(x); }

// Wright (1997), Table A1.
constexpr double WA0 = 7.057924e-4, WA1 = 3.480336e-7, WA2 = -1.112733e-7;
constexpr double WB0 = 5.790749e8, WB1 = 3.516535e6, WB2 = -4.002714e4;
constexpr double WB3 = 2.084372e2, WB4 = 5.944068e5, WB5 = -9.643486e3;
constexpr double WC0 = 1.704853e5, WC1 = 7.904722e2, WC2 = -7.984422e0;
constexpr double WC3 = 5.140652e-2, WC4 = -2.302158e2, WC5 = -3.079464e0;

// Roquet SpV normalisation.
constexpr double kPa2kb = 1.0e-8;    // Pa -> kbar
constexpr double kRdeltaS = 24.0;    // salinity offset, g/kg
constexpr double kR1S0 = 0.875 / 35.16504;
constexpr double kITs = 0.025;       // 1/degC
constexpr double kSrFactor = 35.16504 / 35.0;  // SP -> SR
constexpr double kCp0 = 3991.86795711963;
constexpr double kCtSfac = 0.0248826675584615;

// reference-profile pressure polynomial
constexpr double V00 = -4.4015007269e-05 * kPa2kb;
constexpr double V01 = 6.9232335784e-06 * pw2(kPa2kb);
constexpr double V02 = -7.5004675975e-07 * pw3(kPa2kb);
constexpr double V03 = 1.7009109288e-08 * pw4(kPa2kb);
constexpr double V04 = -1.6884162004e-08 * pw5(kPa2kb);
constexpr double V05 = 1.9613503930e-09 * pw6(kPa2kb);

// SV(zs, zt, zp) = sum SPV_abc zs^a zt^b zp^c
constexpr double S000 = 1.0772899069e-03, S100 = -3.1263658781e-04;
constexpr double S200 = 6.7615860683e-04, S300 = -8.6127884515e-04;
constexpr double S400 = 5.9010812596e-04, S500 = -2.1503943538e-04;
constexpr double S600 = 3.2678954455e-05;
constexpr double S010 = -1.4949652640e-05 * kITs, S110 = 3.1866349188e-05 * kITs;
constexpr double S210 = -3.8070687610e-05 * kITs, S310 = 2.9818473563e-05 * kITs;
constexpr double S410 = -1.0011321965e-05 * kITs, S510 = 1.0751931163e-06 * kITs;
constexpr double S020 = 2.7546851539e-05 * pw2(kITs), S120 = -3.6597334199e-05 * pw2(kITs);
constexpr double S220 = 3.4489154625e-05 * pw2(kITs), S320 = -1.7663254122e-05 * pw2(kITs);
constexpr double S420 = 3.5965131935e-06 * pw2(kITs);
constexpr double S030 = -1.6506828994e-05 * pw3(kITs), S130 = 2.4412359055e-05 * pw3(kITs);
constexpr double S230 = -1.4606740723e-05 * pw3(kITs), S330 = 2.3293406656e-06 * pw3(kITs);
constexpr double S040 = 6.7896174634e-06 * pw4(kITs), S140 = -8.7951832993e-06 * pw4(kITs);
constexpr double S240 = 4.4249040774e-06 * pw4(kITs);
constexpr double S050 = -7.2535743349e-07 * pw5(kITs), S150 = -3.4680559205e-07 * pw5(kITs);
constexpr double S060 = 1.9041365570e-07 * pw6(kITs);
constexpr double S001 = -1.6889436589e-05 * kPa2kb, S101 = 2.1106556158e-05 * kPa2kb;
constexpr double S201 = -2.1322804368e-05 * kPa2kb, S301 = 1.7347655458e-05 * kPa2kb;
constexpr double S401 = -4.3209400767e-06 * kPa2kb;
constexpr double S011 = 1.5355844621e-05 * (kITs * kPa2kb);
constexpr double S111 = 2.0914122241e-06 * (kITs * kPa2kb);
constexpr double S211 = -5.7751479725e-06 * (kITs * kPa2kb);
constexpr double S311 = 1.0767234341e-06 * (kITs * kPa2kb);
constexpr double S021 = -9.6659393016e-06 * (pw2(kITs) * kPa2kb);
constexpr double S121 = -7.0686982208e-07 * (pw2(kITs) * kPa2kb);
constexpr double S221 = 1.4488066593e-06 * (pw2(kITs) * kPa2kb);
constexpr double S031 = 3.1134283336e-06 * (pw3(kITs) * kPa2kb);
constexpr double S131 = 7.9562529879e-08 * (pw3(kITs) * kPa2kb);
constexpr double S041 = -5.6590253863e-07 * (pw4(kITs) * kPa2kb);
constexpr double S002 = 1.0500241168e-06 * pw2(kPa2kb);
constexpr double S102 = 1.9600661704e-06 * pw2(kPa2kb);
constexpr double S202 = -2.1666693382e-06 * pw2(kPa2kb);
constexpr double S012 = -3.8541359685e-06 * (kITs * pw2(kPa2kb));
constexpr double S112 = 1.0157632247e-06 * (kITs * pw2(kPa2kb));
constexpr double S022 = 1.7178343158e-06 * (pw2(kITs) * pw2(kPa2kb));
constexpr double S003 = -4.1503454190e-07 * pw3(kPa2kb);
constexpr double S103 = 3.5627020989e-07 * pw3(kPa2kb);
constexpr double S013 = -1.1293871415e-07 * (kITs * pw3(kPa2kb));

// dSV/dCT = d/dzt of the SV polynomial
constexpr double A000 = S010, A100 = S110, A200 = S210, A300 = S310, A400 = S410, A500 = S510;
constexpr double A010 = 2.0 * S020, A110 = 2.0 * S120, A210 = 2.0 * S220, A310 = 2.0 * S320;
constexpr double A410 = 2.0 * S420;
constexpr double A020 = 3.0 * S030, A120 = 3.0 * S130, A220 = 3.0 * S230, A320 = 3.0 * S330;
constexpr double A030 = 4.0 * S040, A130 = 4.0 * S140, A230 = 4.0 * S240;
constexpr double A040 = 5.0 * S050, A140 = 5.0 * S150, A050 = 6.0 * S060;
constexpr double A001 = S011, A101 = S111, A201 = S211, A301 = S311;
constexpr double A011 = 2.0 * S021, A111 = 2.0 * S121, A211 = 2.0 * S221;
constexpr double A021 = 3.0 * S031, A121 = 3.0 * S131, A031 = 4.0 * S041;
constexpr double A002 = S012, A102 = S112, A012 = 2.0 * S022, A003 = S013;

// dSV/dSA with the 0.5*r1_S0 chain factor folded per term (1/zs applied
// at the call site)
constexpr double B000 = (0.5 * S100) * kR1S0, B100 = S200 * kR1S0;
constexpr double B200 = (1.5 * S300) * kR1S0, B300 = (2.0 * S400) * kR1S0;
constexpr double B400 = (2.5 * S500) * kR1S0, B500 = (3.0 * S600) * kR1S0;
constexpr double B010 = (0.5 * S110) * kR1S0, B110 = S210 * kR1S0;
constexpr double B210 = (1.5 * S310) * kR1S0, B310 = (2.0 * S410) * kR1S0;
constexpr double B410 = (2.5 * S510) * kR1S0;
constexpr double B020 = (0.5 * S120) * kR1S0, B120 = S220 * kR1S0;
constexpr double B220 = (1.5 * S320) * kR1S0, B320 = (2.0 * S420) * kR1S0;
constexpr double B030 = (0.5 * S130) * kR1S0, B130 = S230 * kR1S0;
constexpr double B230 = (1.5 * S330) * kR1S0;
constexpr double B040 = (0.5 * S140) * kR1S0, B140 = S240 * kR1S0;
constexpr double B050 = (0.5 * S150) * kR1S0;
constexpr double B001 = (0.5 * S101) * kR1S0, B101 = S201 * kR1S0;
constexpr double B201 = (1.5 * S301) * kR1S0, B301 = (2.0 * S401) * kR1S0;
constexpr double B011 = (0.5 * S111) * kR1S0, B111 = S211 * kR1S0;
constexpr double B211 = (1.5 * S311) * kR1S0;
constexpr double B021 = (0.5 * S121) * kR1S0, B121 = S221 * kR1S0;
constexpr double B031 = (0.5 * S131) * kR1S0;
constexpr double B002 = (0.5 * S102) * kR1S0, B102 = S202 * kR1S0;
constexpr double B012 = (0.5 * S112) * kR1S0;
constexpr double B003 = (0.5 * S103) * kR1S0;
}  // namespace eosc

template <class R>
__device__ __forceinline__ R rsqrt_exact(R x);
template <>
__device__ __forceinline__ double rsqrt_exact<double>(double x) { return sqrt(x); }
template <>
__device__ __forceinline__ float rsqrt_exact<float>(float x) { return sqrtf(x); }

// roquet_spv_point (redi/ocean_eos.F90:282-448): specific volume and its
// derivatives w.r.t. the model variables (potential temperature, practical
// salinity) at pressure p (Pa).  The body (redi_eos_roquet.inc) is compiled
// twice: as written and without contraction for the plain FP64 form, which
// must round as the reference does; and for the FP32-store form with its
// multiply-adds contracted into FMAs and its divisions by constants taken as
// products with the constants' reciprocals.
template <class R>
__device__ __forceinline__ void roquet_spv_exact(R T_pt, R S_sp, R p, R& sv, R& dsv_dt, R& dsv_ds) {
  const auto cdiv = [](R x, double c) { return x / R(c); };
#include "redi_eos_roquet.inc"
}
template <class R>
__device__ __forceinline__ void roquet_spv_fused(R T_pt, R S_sp, R p, R& sv, R& dsv_dt, R& dsv_ds) {
#pragma clang fp contract(fast)
  const auto cdiv = [](R x, double c) { return x * R(1.0 / c); };
#include "redi_eos_roquet.inc"
}
template <class R>
__device__ __forceinline__ void roquet_spv_point(R T_pt, R S_sp, R p, R& sv, R& dsv_dt,
                                                 R& dsv_ds) {
  if constexpr (std::is_same<R, float>::value)
    roquet_spv_fused<R>(T_pt, S_sp, p, sv, dsv_dt, dsv_ds);
  else
    roquet_spv_exact<R>(T_pt, S_sp, p, sv, dsv_dt, dsv_ds);
}

// Interface density and its T/S sensitivities in the form the neutral sweep
// consumes: dRdT = -rho^2 dSV/dT, dRdS = -rho^2 dSV/dS at (T, S, p)
// (redi/ocean_redi.F90:648-653 over eos_density_point + eos_specvol_derivs).
// For the Roquet variant both reference routines evaluate the same point
// polynomial, so it is evaluated once here.
template <class R>
__device__ __forceinline__ void interface_drho(const EosParams& e, R T, R S, R p, R& dRdT,
                                               R& dRdS) {
  using namespace eosc;
  R rho, dsv_dt, dsv_ds;
  if (e.variant == kEosRoquetSpV) {
    R sv;
    roquet_spv_point<R>(T, S, p, sv, dsv_dt, dsv_ds);
    rho = R(1.0) / sv;
  } else if (e.variant == kEosWright97) {
    const R T_sq = T * T;
    const R a0 = (R(WA0) + R(WA1) * T) + R(WA2) * S;
    const R p_0 = ((((R(WB0) + R(WB1) * T) + R(WB2) * T_sq) + (R(WB3) * T_sq) * T) + R(WB4) * S) +
                  (R(WB5) * S) * T;
    const R lam = ((((R(WC0) + R(WC1) * T) + R(WC2) * T_sq) + (R(WC3) * T_sq) * T) + R(WC4) * S) +
                  (R(WC5) * S) * T;
    const R pp = p + p_0;
    rho = pp / (lam + a0 * pp);
    const R dp0_dt = ((R(WB1) + R(2.0 * WB2) * T) + R(3.0 * WB3) * T_sq) + R(WB5) * S;
    const R dlam_dt = ((R(WC1) + R(2.0 * WC2) * T) + R(3.0 * WC3) * T_sq) + R(WC5) * S;
    const R dp0_ds = R(WB4) + R(WB5) * T;
    const R dlam_ds = R(WC4) + R(WC5) * T;
    const R inv_p = R(1.0) / pp;
    const R inv_p2 = inv_p * inv_p;
    dsv_dt = (R(WA1) + dlam_dt * inv_p) - (lam * dp0_dt) * inv_p2;
    dsv_ds = (R(WA2) + dlam_ds * inv_p) - (lam * dp0_ds) * inv_p2;
  } else {
    const R rho0 = R(e.rho0);
    rho = (rho0 + R(e.beta_S) * (S - R(e.S_ref))) - R(e.alpha_T) * (T - R(e.T_ref));
    dsv_dt = R(e.alpha_T) / (rho0 * rho0);
    dsv_ds = -(R(e.beta_S) / (rho0 * rho0));
  }
  dRdT = -((rho * rho) * dsv_dt);
  dRdS = -((rho * rho) * dsv_ds);
}

}  // namespace redi
