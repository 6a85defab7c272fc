// redi_column.hpp -- per-column reconstruction and neutral-sweep device code.
//
// The reference reconstructs columns into NZ_STACK_MAX-sized per-thread
// arrays (redi/ocean_redi.F90:231-254, 609-654, 852-891), which on CDNA
// becomes scratch spill traffic.  Here a column is streamed top-down once
// with a four-layer register window: the PPM edge at interface k only needs
// layers k-2..k+1 and the PLM slopes of layers k-1 and k, so each layer is
// loaded exactly once (coalesced along i, since lanes are adjacent columns)
// and every interface value is emitted as soon as it is known.
//
// The arithmetic of every helper below is the reference's, operation by
// operation (see the cited lines), so the FP64 instantiation compiled without
// FP contraction reproduces the reference bits.
#pragma once

#include <hip/hip_runtime.h>

namespace redi {

constexpr double kGravity = 9.80665;   // redi/constants.F90:19
constexpr double kHDivEps = 1.0e-20;   // redi/constants.F90:21
constexpr double kHNeglect = 1.0e-30;  // redi/ocean_redi.F90:242

template <class R>
__device__ __forceinline__ R f_sign(R a, R b) {  // Fortran SIGN(a, b)
  return copysign(fabs(a), b);
}
template <class R>
__device__ __forceinline__ R fmin_sel(R a, R b) { return a < b ? a : b; }
// The reference's sign test redi_signum(1,a) * redi_signum(1,b) <= 0
// (redi_signum(1,x) is 0 at x == 0, else 1 with the sign of x), without the
// products: true when either value is zero or their sign bits differ (a NaN
// counts with its sign bit, as copysign gives it), so the same decisions.
template <class R>
__device__ __forceinline__ bool sgn_prod_le0(R a, R b) {
  return a == R(0) || b == R(0) || (__builtin_signbit(a) != __builtin_signbit(b));
}
template <class R>
__device__ __forceinline__ R fmax_sel(R a, R b) { return a > b ? a : b; }

// 1 / x: the IEEE quotient in FP64 (the plain form rounds as the
// reference); in FP32 the hardware reciprocal (v_rcp_f32, within 1 ulp),
// which the FP32-store form's reconstruction weights use.
template <class R>
__device__ __forceinline__ R recip(R x) { return R(1) / x; }
template <>
__device__ __forceinline__ float recip<float>(float x) { return __builtin_amdgcn_rcpf(x); }

// The PLM slope and the PPM edge are linear in the tracer values with
// weights that depend on the layer thicknesses alone.  Every tracer sharing a
// thickness column shares those weights, so they are formed once per layer
// (plm_weights, ppm_weights) and applied per tracer (plm_slope_w,
// ppm_edge_w).  Each weight is the reference's own subexpression, evaluated
// in the reference's order, so weights + application give the same bits as
// the one-shot helpers below them.

// thickness part of redi_fv_diff (redi/ocean_redi.F90:135-150) and of the
// PLM guard (:174-187)
template <class R>
struct PlmW {
  R wa, wb, wc;  // (2 hkm1 + hk) / (hkp1 + hk), (2 hkp1 + hk) / (hkm1 + hk), hk / h_sum
  bool flat;     // the slope is 0 whatever the tracer values
};
template <class R>
__device__ __forceinline__ PlmW<R> plm_weights(R hkm1, R hk, R hkp1) {
  R h_sum = (hkm1 + hkp1) + hk;
  if (h_sum != R(0)) h_sum = recip<R>(h_sum);
  R hm = hkm1 + hk;
  if (hm != R(0)) hm = recip<R>(hm);
  R hp = hkp1 + hk;
  if (hp != R(0)) hp = recip<R>(hp);
  return PlmW<R>{(R(2) * hkm1 + hk) * hp, (R(2) * hkp1 + hk) * hm, hk * h_sum,
                 !((hkp1 + hk) * (hkm1 + hk) > R(0))};
}

// redi_fv_diff (redi/ocean_redi.F90:135-150) from its weights
template <class R>
__device__ __forceinline__ R fv_diff_w(const PlmW<R>& w, R skm1, R sk, R skp1) {
  const R a = w.wa * (skp1 - sk);
  const R b = w.wb * (sk - skm1);
  return w.wc * (a + b);
}

// one interior entry of redi_plm_diff (redi/ocean_redi.F90:174-187)
template <class R>
__device__ __forceinline__ R plm_slope_w(const PlmW<R>& w, R sm1, R s0, R sp1) {
  if (w.flat) return R(0);
  const R dc = fv_diff_w<R>(w, sm1, s0, sp1);
  const R dl = R(2) * (s0 - sm1);
  const R dr = R(2) * (sp1 - s0);
  if (sgn_prod_le0<R>(dl, dr)) return R(0);
  const R m = fmin_sel<R>(fmin_sel<R>(fabs(dl), fabs(dc)), fabs(dr));
  return f_sign<R>(m, dc);
}

template <class R>
__device__ __forceinline__ R fv_diff(R hkm1, R hk, R hkp1, R skm1, R sk, R skp1) {
  return fv_diff_w<R>(plm_weights<R>(hkm1, hk, hkp1), skm1, sk, skp1);
}
template <class R>
__device__ __forceinline__ R plm_slope(R hm1, R h0, R hp1, R sm1, R s0, R sp1) {
  return plm_slope_w<R>(plm_weights<R>(hm1, h0, hp1), sm1, s0, sp1);
}

// thickness part of redi_ppm_edge (redi/ocean_redi.F90:194-225)
template <class R>
struct PpmW {
  R we;          // hk / (hk + hkp1) or hkp1 / (hk + hkp1), by `lower`
  R f1, f2, f3, f4;
  bool mean;     // hk + hkp1 <= 0: the edge is the mean of the two values
  bool lower;    // hk < hkp1: the edge is built from ak
};
template <class R>
__device__ __forceinline__ PpmW<R> ppm_weights(R hkm1, R hk, R hkp1, R hkp2) {
  PpmW<R> w;
  R r = hk + hkp1;
  w.mean = r <= R(0);
  r = recip<R>(r);
  w.lower = hk < hkp1;
  w.we = w.lower ? hk * r : hkp1 * r;
  const R r2a = recip<R>((R(2) * hk + hkp1) + R(kHNeglect));
  const R r2b = recip<R>((hk + R(2) * hkp1) + R(kHNeglect));
  w.f1 = recip<R>((hk + hkp1) + (hkm1 + hkp2));
  w.f2 = ((R(2) * (hkp1 * hk)) * r) * (((hkm1 + hk) * r2a) - ((hkp2 + hkp1) * r2b));
  w.f3 = (hk * (hkm1 + hk)) * r2a;
  w.f4 = (hkp1 * (hkp1 + hkp2)) * r2b;
  return w;
}

// redi_ppm_edge (redi/ocean_redi.F90:194-225) from its weights
template <class R>
__device__ __forceinline__ R ppm_edge_w(const PpmW<R>& w, R ak, R akp1, R pk, R pkp1) {
  if (w.mean) return R(0.5) * (ak + akp1);
  const R e = w.lower ? ak + w.we * (akp1 - ak) : akp1 + w.we * (ak - akp1);
  return e + w.f1 * (w.f2 * (akp1 - ak) - (w.f3 * pkp1 - w.f4 * pk));
}

template <class R>
__device__ __forceinline__ R ppm_edge(R hkm1, R hk, R hkp1, R hkp2, R ak, R akp1, R pk, R pkp1) {
  return ppm_edge_w<R>(ppm_weights<R>(hkm1, hk, hkp1, hkp2), ak, akp1, pk, pkp1);
}

// Streaming form of redi_interface_scalar (redi/ocean_redi.F90:231-254) for
// NS scalars sharing one thickness column.  `load(kt, h, s[])` returns the
// top-down layer kt (0-based; the caller flips the native bottom-up index);
// `emit(k, e[], s_above[])` receives interface k = 0..nz (0 = surface) with
// the edge values and, for k >= 1, the layer-average of layer k-1 above it.
template <class R, int NS, class Load, class Emit>
__device__ __forceinline__ void stream_edges(int nz, Load load, Emit emit) {
  R h0, h1, h2, s0[NS], s1[NS], s2[NS];
  load(0, h0, s0);
  load(nz > 1 ? 1 : 0, h1, s1);
  load(nz > 2 ? 2 : (nz - 1), h2, s2);
  // window at interface k: hw0..hw3 = h[k-2], h[k-1], h[k], h[k+1] (clamped)
  R hw0 = h0, hw1 = h0, hw2 = h1, hw3 = h2;
  R sw0[NS], sw1[NS], sw2[NS], dprev[NS], e[NS];
#pragma unroll
  for (int q = 0; q < NS; ++q) {
    sw0[q] = s0[q]; sw1[q] = s1[q]; sw2[q] = s2[q]; dprev[q] = R(0);
    e[q] = s0[q] - R(0.5) * R(0);  // edge(1) = tr(1) - 0.5*diff(1), diff(1) = 0
  }
  emit(0, e, sw0);
  for (int k = 1; k < nz; ++k) {
    R dcur[NS];
    const PlmW<R> pw = plm_weights<R>(hw1, hw2, hw3);
    const PpmW<R> ew = ppm_weights<R>(hw0, hw1, hw2, hw3);
#pragma unroll
    for (int q = 0; q < NS; ++q) {
      dcur[q] = (k <= nz - 2) ? plm_slope_w<R>(pw, sw0[q], sw1[q], sw2[q]) : R(0);
      e[q] = ppm_edge_w<R>(ew, sw0[q], sw1[q], dprev[q], dcur[q]);
    }
    emit(k, e, sw0);
    R hn, sn[NS];
    const int kn = (k + 2 < nz) ? (k + 2) : (nz - 1);
    load(kn, hn, sn);
    hw0 = hw1; hw1 = hw2; hw2 = hw3; hw3 = hn;
#pragma unroll
    for (int q = 0; q < NS; ++q) {
      sw0[q] = sw1[q]; sw1[q] = sw2[q]; sw2[q] = sn[q]; dprev[q] = dcur[q];
    }
  }
  // after the loop sw0 holds layer nz-1 for every nz >= 1
#pragma unroll
  for (int q = 0; q < NS; ++q) e[q] = sw0[q] + R(0.5) * R(0);  // diff(nk) = 0
  emit(nz, e, sw0);
}

// stream_edges for the FP32-store form: FP64 layer values (`load` returns
// them), PLM slopes and PPM edges computed in FP32 on differences about the
// FP64 layer means -- a slope about its own layer's mean, an edge about the
// mean of the layer above it -- and each edge returned as that mean plus
// its FP32 offset.  Shifting a column's values by a constant shifts its
// edges by the same constant and leaves its slopes unchanged, so this is the
// FP32 evaluation of the same formulas (as in the tiled reconstruction).
template <int NS, class Load, class Emit>
__device__ __forceinline__ void stream_edges_off(int nz, Load load, Emit emit) {
  double h0, h1, h2, s0[NS], s1[NS], s2[NS];
  load(0, h0, s0);
  load(nz > 1 ? 1 : 0, h1, s1);
  load(nz > 2 ? 2 : (nz - 1), h2, s2);
  float hw0 = (float)h0, hw1 = (float)h0, hw2 = (float)h1, hw3 = (float)h2;
  double sw0[NS], sw1[NS], sw2[NS], e[NS];
  float dprev[NS];
#pragma unroll
  for (int q = 0; q < NS; ++q) {
    sw0[q] = s0[q]; sw1[q] = s1[q]; sw2[q] = s2[q]; dprev[q] = 0.f;
    e[q] = s0[q];
  }
  emit(0, e, sw0);
  for (int k = 1; k < nz; ++k) {
    float dcur[NS];
    const PlmW<float> pw = plm_weights<float>(hw1, hw2, hw3);
    const PpmW<float> ew = ppm_weights<float>(hw0, hw1, hw2, hw3);
#pragma unroll
    for (int q = 0; q < NS; ++q) {
      dcur[q] = (k <= nz - 2)
                    ? plm_slope_w<float>(pw, (float)(sw0[q] - sw1[q]), 0.f, (float)(sw2[q] - sw1[q]))
                    : 0.f;
      e[q] = sw0[q] + (double)ppm_edge_w<float>(ew, 0.f, (float)(sw1[q] - sw0[q]), dprev[q], dcur[q]);
    }
    emit(k, e, sw0);
    double hn, sn[NS];
    const int kn = (k + 2 < nz) ? (k + 2) : (nz - 1);
    load(kn, hn, sn);
    hw0 = hw1; hw1 = hw2; hw2 = hw3; hw3 = (float)hn;
#pragma unroll
    for (int q = 0; q < NS; ++q) {
      sw0[q] = sw1[q]; sw1[q] = sw2[q]; sw2[q] = sn[q]; dprev[q] = dcur[q];
    }
  }
#pragma unroll
  for (int q = 0; q < NS; ++q) e[q] = sw0[q];
  emit(nz, e, sw0);
}

// PPM limiter of redi_tracer_column (redi/ocean_redi.F90:876-890): limits
// the edges al / ar of a layer of mean tl in place.  Returns the branch
// taken: 0 edges kept, 1 both flattened to the mean, 2 the top edge rebuilt
// from the mean and the bottom one, 3 the bottom edge rebuilt from the mean
// and the top one.
template <class R>
__device__ __forceinline__ int ppm_limit(R tl, R& al, R& ar) {
  if (sgn_prod_le0<R>(ar - tl, tl - al)) {
    al = tl;
    ar = tl;
    return 1;
  } else if (f_sign<R>(R(3), ar - al) * ((tl - al) + (tl - ar)) > fabs(ar - al)) {
    al = tl + R(2) * (tl - ar);
    return 2;
  } else if (f_sign<R>(R(3), ar - al) * ((tl - al) + (tl - ar)) < -fabs(ar - al)) {
    ar = tl + R(2) * (tl - al);
    return 3;
  }
  return 0;
}

// redi_interpolate_position (redi/ocean_redi.F90:260-285)
template <class R>
__device__ __forceinline__ R interpolate_position(R dRhoNeg, R Pneg, R dRhoPos, R Ppos) {
  if ((Ppos > Pneg) && (dRhoPos - dRhoNeg >= R(0))) {
    if (dRhoPos - dRhoNeg > R(0))
      return fmin_sel<R>(R(1), fmax_sel<R>(R(0), -dRhoNeg / (dRhoPos - dRhoNeg)));
    if (dRhoNeg > R(0)) return R(0);
    if (dRhoNeg < R(0)) return R(1);
    return R(0.5);
  }
  return R(0.5);
}

// redi_ppm_ave (redi/ocean_redi.F90:833-850)
template <class R>
__device__ __forceinline__ R ppm_ave(R xL, R xR, R aL, R aR, R aMean) {
  const R dx = xR - xL;
  const R xave = R(0.5) * (xR + xL);
  const R a6o3 = R(2) * aMean - (aL + aR);
  const R a6 = R(3) * a6o3;
  if (dx > R(0) && dx <= R(1))
    return (aL + xave * ((aR - aL) + a6)) - a6o3 * ((xR * xR + xR * xL) + xL * xL);
  return (aL + (aR - aL) * xR) + (a6 * xR) * (R(1) - xR);
}

}  // namespace redi
