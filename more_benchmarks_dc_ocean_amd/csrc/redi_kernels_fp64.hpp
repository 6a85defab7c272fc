// redi_kernels_fp64.hpp -- the plain FP64 Redi step as four gfx950 kernels.
//
//   k_column_state   Phase A, once per column: top-down interface pressure,
//                    PPM T/S edges and dR/dT, dR/dS from the EOS.
//   k_neutral_sweep  Phase A, once per wet face: the continuous neutral
//                    surface sweep over the two adjacent columns, writing
//                    PoL/PoR/KoL/KoR/hEff in the reference layout.
//   k_tracer_recon   Phase B, once per column and tracer: layer averages,
//                    PPM edges and limited left/right edge values.
//   k_apply_flux     Phase B, once per cell and tracer: the four bounding
//                    faces' sublayer fluxes accumulated into the cell's own
//                    layers (W, E, S, N order), then hTr += dTr / areaT.
//
// Memory layout (all device, FP64 unless noted):
//   3-D cell fields (h, hTr):           [k][j][i], k native bottom-up
//   column interface scratch (K1/K3):   [kt][j][i], kt top-down interface
//   face coefficients (Po/Ko/hEff):     [s][j][i] of the face grid (Ko int32)
// i.e. exactly the reference's Fortran column-major arrays, so lanes of a
// wavefront (consecutive i) always touch consecutive addresses for a given
// layer / interface / surface.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "redi_column.hpp"
#include "redi_eos.hpp"

namespace redi {

// ---------------------------------------------------------------------------
// K1: column interface state (redi_build_column, redi/ocean_redi.F90:609-654)
// ---------------------------------------------------------------------------
struct ColumnState {
  double* P;     // interface pressure, [nz+1][N2]
  double* T;     // interface T
  double* S;     // interface S
  double* dRdT;  // -rho^2 dSV/dT
  double* dRdS;  // -rho^2 dSV/dS
};

__global__ void __launch_bounds__(256) k_column_state(int N2, int nz, const double* __restrict__ h,
                                                      const double* __restrict__ thtr,
                                                      const double* __restrict__ shtr, EosParams eos,
                                                      ColumnState cs) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= N2) return;
  const size_t n2 = (size_t)N2;
  const double g_rho0 = kGravity * eos.rho0;
  double P = 0.0;     // Pint(1) = 0
  double h_above = 0.0;
  auto load = [&](int kt, double& hh, double* s) {
    const size_t o = (size_t)(nz - 1 - kt) * n2 + c;  // top-down -> native
    hh = h[o];
    const double he = hh > kHDivEps ? hh : kHDivEps;
    s[0] = thtr[o] / he;
    s[1] = shtr[o] / he;
  };
  // The thickness of the layer above interface k is needed for Pint(k+1);
  // it is re-read (cached) rather than carried through the edge window.
  auto emit = [&](int k, const double* e, const double* /*s_above*/) {
    if (k > 0) {
      h_above = h[(size_t)(nz - k) * n2 + c];  // top-down layer k-1
      P = P + g_rho0 * h_above;
    }
    double dRdT, dRdS;
    interface_drho<double>(eos, e[0], e[1], P, dRdT, dRdS);
    const size_t o = (size_t)k * n2 + c;
    cs.P[o] = P;
    cs.T[o] = e[0];
    cs.S[o] = e[1];
    cs.dRdT[o] = dRdT;
    cs.dRdS[o] = dRdS;
  };
  stream_edges<double, 2>(nz, load, emit);
}

// ---------------------------------------------------------------------------
// K2: continuous neutral-surface sweep
// (redi_neutral_positions_continuous, redi/ocean_redi.F90:307-464, run by
//  redi_face_coeffs / redi_calc_coeffs_x/_y, :656-812)
// ---------------------------------------------------------------------------
struct FaceCoeffs {
  double* PoL;
  double* PoR;
  int32_t* KoL;
  int32_t* KoR;
  double* hEff;
};

struct ColView {  // one column of the K1 scratch; k is the 1-based interface
  const double *P, *T, *S, *dT, *dS;
  size_t c, n2;
  __device__ __forceinline__ size_t at(int k) const { return (size_t)(k - 1) * n2 + c; }
};

// secant density difference rho(X, a) - rho(Y, b)
__device__ __forceinline__ double sec_drho(const ColView& X, int a, const ColView& Y, int b) {
  const size_t xa = X.at(a), yb = Y.at(b);
  return 0.5 * ((X.dT[xa] + Y.dT[yb]) * (X.T[xa] - Y.T[yb]) +
                (X.dS[xa] + Y.dS[yb]) * (X.S[xa] - Y.S[yb]));
}

// Face grid: faces (i, j) for i in [i0, i0+nfi), j in [j0, j0+nfj) of an
// (nfa x nfb) face array; the left column is (i-di, j-dj) and the right
// column (i, j) of the (nx x ny) cell grid.
__global__ void __launch_bounds__(256) k_neutral_sweep(
    int nx, int N2, int nz, int nfa, int nfb, int i0, int j0, int nfi, int nfj, int di, int dj,
    const double* __restrict__ wet, ColumnState cs, FaceCoeffs fc, double pa_to_h) {
  const long f = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= (long)nfi * nfj) return;
  const int i = i0 + (int)(f % nfi), j = j0 + (int)(f / nfi);
  const size_t fo = (size_t)i + (size_t)nfa * j;
  const size_t fs = (size_t)nfa * nfb;
  const int nk = nz, ns = 2 * nz + 2;
  if (!(wet[fo] > 0.0)) {
    for (int s = 0; s < ns; ++s) {
      fc.PoL[fo + s * fs] = 0.0;
      fc.PoR[fo + s * fs] = 0.0;
      fc.KoL[fo + s * fs] = 1;
      fc.KoR[fo + s * fs] = 1;
    }
    for (int s = 0; s < ns - 1; ++s) fc.hEff[fo + s * fs] = 0.0;
    return;
  }
  const size_t n2 = (size_t)N2;
  const ColView L{cs.P, cs.T, cs.S, cs.dRdT, cs.dRdS, (size_t)(i - di) + (size_t)nx * (j - dj), n2};
  const ColView Rc{cs.P, cs.T, cs.S, cs.dRdT, cs.dRdS, (size_t)i + (size_t)nx * j, n2};

  int kr = 1, kl = 1, lastK_r = 1, lastK_l = 1;
  double lastP_r = 0.0, lastP_l = 0.0, lastPos_l = 0.0, lastPos_r = 0.0;
  bool reached_bottom = false, search_l = false;
  for (int ks = 1; ks <= ns; ++ks) {
    int klm1 = kl - 1 > 1 ? kl - 1 : 1;
    int krm1 = kr - 1 > 1 ? kr - 1 : 1;
    if (!reached_bottom) {
      const double dRho = sec_drho(Rc, kr, L, kl);
      if (dRho < 0.0)
        search_l = true;
      else if (dRho > 0.0)
        search_l = false;
      else if (kl + kr == 2)
        search_l = true;
      else
        search_l = !search_l;
    }
    double PoL, PoR;
    int KoL, KoR;
    if (search_l) {
      const double top = sec_drho(L, klm1, Rc, kr);
      const double bot = sec_drho(L, klm1 + 1, Rc, kr);
      if (top > 0.0 || kr + kl == 2)
        PoL = 0.0;
      else if (top >= bot)
        PoL = 1.0;
      else
        PoL = interpolate_position<double>(top, L.P[L.at(klm1)], bot, L.P[L.at(klm1 + 1)]);
      if (PoL >= 1.0 && klm1 < nk) {
        klm1 += 1;
        PoL = PoL - 1.0;
      }
      if ((double)(klm1 - lastK_l) + (PoL - lastP_l) < 0.0) {
        PoL = lastP_l;
        klm1 = lastK_l;
      }
      KoL = klm1;
      if (kr <= nk) {
        PoR = 0.0; KoR = kr; kr += 1;
      } else {
        PoR = 1.0; KoR = nk;
        reached_bottom = true; search_l = false;
      }
    } else {
      const double top = sec_drho(Rc, krm1, L, kl);
      const double bot = sec_drho(Rc, krm1 + 1, L, kl);
      if (top >= 0.0 || kr + kl == 2)
        PoR = 0.0;
      else if (top >= bot)
        PoR = 1.0;
      else
        PoR = interpolate_position<double>(top, Rc.P[Rc.at(krm1)], bot, Rc.P[Rc.at(krm1 + 1)]);
      if (PoR >= 1.0 && krm1 < nk) {
        krm1 += 1;
        PoR = PoR - 1.0;
      }
      if ((double)(krm1 - lastK_r) + (PoR - lastP_r) < 0.0) {
        PoR = lastP_r;
        krm1 = lastK_r;
      }
      KoR = krm1;
      if (kl <= nk) {
        PoL = 0.0; KoL = kl; kl += 1;
      } else {
        PoL = 1.0; KoL = nk;
        reached_bottom = true; search_l = true;
      }
    }
    lastK_l = KoL; lastP_l = PoL; lastK_r = KoR; lastP_r = PoR;
    const size_t so = fo + (size_t)(ks - 1) * fs;
    fc.PoL[so] = PoL;
    fc.PoR[so] = PoR;
    fc.KoL[so] = KoL;
    fc.KoR[so] = KoR;
    // redi_absolute_position (:289-298) of this surface in both columns
    const double pl0 = L.P[L.at(KoL)], pr0 = Rc.P[Rc.at(KoR)];
    const double posL = pl0 + PoL * (L.P[L.at(KoL + 1)] - pl0);
    const double posR = pr0 + PoR * (Rc.P[Rc.at(KoR + 1)] - pr0);
    if (ks > 1) {
      const double hL = posL - lastPos_l;
      const double hR = posR - lastPos_r;
      double he = 0.0;
      if (hL + hR > 0.0) he = ((2.0 * hL) * hR) / (hL + hR);
      fc.hEff[fo + (size_t)(ks - 2) * fs] = he * pa_to_h;  // Pa -> m (:688-691)
    }
    lastPos_l = posL;
    lastPos_r = posR;
  }
}

// ---------------------------------------------------------------------------
// K3: tracer column reconstruction (redi_tracer_column, :852-891)
// ---------------------------------------------------------------------------
struct TracerRecon {
  double* Tl;  // layer average, [nz][N2] top-down
  double* Ti;  // interface edge, [nz+1][N2]
  double* aL;  // limited upper edge, [nz][N2]
  double* aR;  // limited lower edge, [nz][N2]
};

__global__ void __launch_bounds__(256) k_tracer_recon(int N2, int nz, const double* __restrict__ h,
                                                      const double* __restrict__ htr,
                                                      TracerRecon tr) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= N2) return;
  const size_t n2 = (size_t)N2;
  auto load = [&](int kt, double& hh, double* s) {
    const size_t o = (size_t)(nz - 1 - kt) * n2 + c;
    hh = h[o];
    const double he = hh > kHDivEps ? hh : kHDivEps;
    s[0] = htr[o] / he;
  };
  double e_prev = 0.0;
  auto emit = [&](int k, const double* e, const double* s_above) {
    tr.Ti[(size_t)k * n2 + c] = e[0];
    if (k > 0) {
      const double tl = s_above[0];
      double al = e_prev, ar = e[0];
      ppm_limit<double>(tl, al, ar);
      const size_t o = (size_t)(k - 1) * n2 + c;
      tr.Tl[o] = tl;
      tr.aL[o] = al;
      tr.aR[o] = ar;
    }
    e_prev = e[0];
  };
  stream_edges<double, 1>(nz, load, emit);
}

// ---------------------------------------------------------------------------
// K4: sublayer fluxes and divergence (redi_apply_flux_impl + redi_face_flux
// + redi_sublayer_dT, :893-925, :1037-1181)
// ---------------------------------------------------------------------------
struct FluxArgs {
  int nx, ny, nz;
  double dt;
  int wuf_w, wuf_e, wvf_s, wvf_n;  // 1-based wall faces (:1139-1142)
  int wall_w, wall_e, wall_s, wall_n;
  const double *khtr_u, *khtr_v, *dy_cu, *dx_cv, *idxCu, *idyCv, *areaT;
  FaceCoeffs u, v;
};

struct ReconView {
  const double *Tl, *Ti, *aL, *aR;
  size_t c, n2;
};

// redi_sublayer_dT with 1-based top-down layer indices
__device__ __forceinline__ double sublayer_dT(int klt, int klb, int krt, int krb, double PoLt,
                                              double PoLb, double PoRt, double PoRb,
                                              const ReconView& L, const ReconView& R) {
  const size_t n2 = L.n2;
  const double tlt = (1.0 - PoLt) * L.Ti[(size_t)(klt - 1) * n2 + L.c] + PoLt * L.Ti[(size_t)klt * n2 + L.c];
  const double tlb = (1.0 - PoLb) * L.Ti[(size_t)(klb - 1) * n2 + L.c] + PoLb * L.Ti[(size_t)klb * n2 + L.c];
  const double trt = (1.0 - PoRt) * R.Ti[(size_t)(krt - 1) * n2 + R.c] + PoRt * R.Ti[(size_t)krt * n2 + R.c];
  const double trb = (1.0 - PoRb) * R.Ti[(size_t)(krb - 1) * n2 + R.c] + PoRb * R.Ti[(size_t)krb * n2 + R.c];
  const size_t ol = (size_t)(klt - 1) * n2 + L.c, orr = (size_t)(krt - 1) * n2 + R.c;
  const double tlay = ppm_ave<double>(PoLt, PoLb + (double)(klb - klt), L.aL[ol], L.aR[ol], L.Tl[ol]);
  const double trlay = ppm_ave<double>(PoRt, PoRb + (double)(krb - krt), R.aL[orr], R.aR[orr], R.Tl[orr]);
  const double dT_top = trt - tlt;
  const double dT_bot = trb - tlb;
  const double dT_ave = 0.5 * (dT_top + dT_bot);
  const double dT_layer = trlay - tlay;
  if (signum1<double>(dT_top) * signum1<double>(dT_bot) <= 0.0 ||
      signum1<double>(dT_ave) * signum1<double>(dT_layer) <= 0.0)
    return 0.0;
  return dT_layer;
}

// One face's ns-1 sublayers into the LDS accumulator of this cell.
__device__ __forceinline__ void face_flux(const FaceCoeffs& fc, size_t fo, size_t fs, int nz,
                                          const ReconView& L, const ReconView& R, double coef,
                                          bool centre_is_left, double* acc, int stride) {
  const int ns = 2 * nz + 2;
  for (int s = 0; s < ns - 1; ++s) {
    const size_t a = fo + (size_t)s * fs;
    const double he = fc.hEff[a];
    if (he != 0.0) {
      const size_t b = a + fs;
      const int kla = fc.KoL[a], kra = fc.KoR[a];
      const double d = sublayer_dT(kla, fc.KoL[b], kra, fc.KoR[b], fc.PoL[a], fc.PoL[b], fc.PoR[a],
                                   fc.PoR[b], L, R);
      const double flx = (d * he) * coef;
      if (centre_is_left) {
        double* p = acc + (size_t)(nz - kla) * stride;
        *p = *p + flx;
      } else {
        double* p = acc + (size_t)(nz - kra) * stride;
        *p = *p - flx;
      }
    }
  }
}

// blockDim.x threads, dynamic LDS = nz * blockDim.x doubles (dTr per thread,
// layer-major so simultaneous accesses to one layer hit distinct banks).
__global__ void __launch_bounds__(64) k_apply_flux(FluxArgs a, TracerRecon rc,
                                                   const double* __restrict__ /*h*/,
                                                   double* __restrict__ htr) {
  extern __shared__ double lds[];
  const int N2 = a.nx * a.ny;
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  const int stride = blockDim.x;
  double* acc = lds + threadIdx.x;
  const int nz = a.nz;
  for (int k = 0; k < nz; ++k) acc[(size_t)k * stride] = 0.0;
  if (c >= N2) return;
  const int i = c % a.nx + 1, j = c / a.nx + 1;  // 1-based
  const size_t n2 = (size_t)N2;
  const ReconView C{rc.Tl, rc.Ti, rc.aL, rc.aR, (size_t)c, n2};
  const size_t ufs = (size_t)(a.nx + 1) * a.ny, vfs = (size_t)a.nx * (a.ny + 1);
  // WEST u-face (face i): this cell is the right column
  if (i >= 2 && !((a.wall_w && i == a.wuf_w) || (a.wall_e && i == a.wuf_e))) {
    const size_t fo = (size_t)(i - 1) + (size_t)(a.nx + 1) * (j - 1);
    const ReconView N{rc.Tl, rc.Ti, rc.aL, rc.aR, (size_t)c - 1, n2};
    const double coef = ((a.dt * a.khtr_u[fo]) * a.dy_cu[fo]) * a.idxCu[fo];
    face_flux(a.u, fo, ufs, nz, N, C, coef, false, acc, stride);
  }
  // EAST u-face (face i+1): this cell is the left column
  if (i <= a.nx - 1 && !((a.wall_w && i + 1 == a.wuf_w) || (a.wall_e && i + 1 == a.wuf_e))) {
    const size_t fo = (size_t)i + (size_t)(a.nx + 1) * (j - 1);
    const ReconView N{rc.Tl, rc.Ti, rc.aL, rc.aR, (size_t)c + 1, n2};
    const double coef = ((a.dt * a.khtr_u[fo]) * a.dy_cu[fo]) * a.idxCu[fo];
    face_flux(a.u, fo, ufs, nz, C, N, coef, true, acc, stride);
  }
  // SOUTH v-face (face j): this cell is the right (north) column
  if (j >= 2 && !((a.wall_s && j == a.wvf_s) || (a.wall_n && j == a.wvf_n))) {
    const size_t fo = (size_t)(i - 1) + (size_t)a.nx * (j - 1);
    const ReconView N{rc.Tl, rc.Ti, rc.aL, rc.aR, (size_t)c - a.nx, n2};
    const double coef = ((a.dt * a.khtr_v[fo]) * a.dx_cv[fo]) * a.idyCv[fo];
    face_flux(a.v, fo, vfs, nz, N, C, coef, false, acc, stride);
  }
  // NORTH v-face (face j+1): this cell is the left (south) column
  if (j <= a.ny - 1 && !((a.wall_s && j + 1 == a.wvf_s) || (a.wall_n && j + 1 == a.wvf_n))) {
    const size_t fo = (size_t)(i - 1) + (size_t)a.nx * j;
    const ReconView N{rc.Tl, rc.Ti, rc.aL, rc.aR, (size_t)c + a.nx, n2};
    const double coef = ((a.dt * a.khtr_v[fo]) * a.dx_cv[fo]) * a.idyCv[fo];
    face_flux(a.v, fo, vfs, nz, C, N, coef, true, acc, stride);
  }
  const double iaij = 1.0 / a.areaT[c];
  for (int k = 0; k < nz; ++k) {
    const size_t o = (size_t)k * n2 + c;
    htr[o] = htr[o] + acc[(size_t)k * stride] * iaij;
  }
}

// khtr_u/v <- khtr_*_ext (redi_face_copy, :1015-1024) or the scalar khtr
// (redi_face_const, :1026-1035); one launch covers both face grids.
__global__ void k_face_khtr(int nu, int nv, const double* __restrict__ ext_u,
                            const double* __restrict__ ext_v, double khtr, double* __restrict__ ku,
                            double* __restrict__ kv) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < nu) {
    ku[t] = ext_u ? ext_u[t] : khtr;
  } else if (t < nu + nv) {
    kv[t - nu] = ext_v ? ext_v[t - nu] : khtr;
  }
}

// ocean_redi_init fill (:493-502): Po = hEff = 0, Ko = 1
__global__ void k_fill_coeffs(size_t n_po, size_t n_he, FaceCoeffs fc) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n_po) {
    fc.PoL[t] = 0.0;
    fc.PoR[t] = 0.0;
    fc.KoL[t] = 1;
    fc.KoR[t] = 1;
  }
  if (t < n_he) fc.hEff[t] = 0.0;
}

}  // namespace redi
