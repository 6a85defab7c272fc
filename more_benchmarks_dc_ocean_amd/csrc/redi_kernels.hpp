// redi_kernels.hpp -- the Redi step as gfx950 kernels, in two precision forms.
//
//   k_column_state        Phase A, once per column: top-down interface
//                         pressure, PPM T/S edges and dR/dT, dR/dS (EOS).
//   k_neutral_sweep       Phase A, once per wet face: the continuous neutral
//                         surface sweep over the two adjacent columns.
//   k_tracer_recon_tiled  Phase B, per tracer group: layer averages, PPM edges
//                         and limited edge values of every column.
//   k_face_flux           Phase B, once per face for up to NT tracers: the flux
//                         of every contributing sublayer and its target layers.
//   k_cell_flux           Phase B, once per cell: the four bounding faces'
//                         sublayer fluxes added into the cell's layers in the
//                         reference's order, then hTr += dTr / areaT.
//   k_export_coeffs       face records -> the reference's Po/Ko/hEff arrays.
//
// Every kernel is a template on a precision form F:
//   Fp64  the plain form: fields, records and arithmetic in FP64, in the
//         reference's rounding order, so results match the reference bits;
//   Fp32  the FP32-store form: the fields h / hTr, the face records, the
//         per-step reconstruction and the sublayer fluxes are stored in
//         32-bit words (FP32 values, int16 layer indices, positions as
//         32-bit fractions; the reconstruction as an FP64 layer mean plus
//         FP32 offsets of its edge values, so differences between columns
//         keep FP64 accuracy).  The arithmetic stays FP64: every difference
//         of nearly equal values -- the T/S edges, EOS derivatives and
//         density differences that place the neutral surfaces, the tracer
//         differences across a face, the tendency accumulation -- is
//         formed in FP64 from the stored values.
// Both forms come from one source: records convert through get() / put(),
// which are identities for Fp64.
//
// Device memory layout (R = the form's storage type):
//   cell fields h, hTr          [k][j][i] (reference layout, k native bottom-up)
//   column state (K1 -> K2)     FP64 in both forms, interface-major within
//                               each row: ColRec / P of column (i, j) at
//                               interface kt (top-down, 0..nz) at
//                               (j*(nz+1) + kt)*nx + i
//   face coefficients (K2->K4)  FaceRec per face and neutral surface s,
//                               surface-major: (face, s) at s*Nf' + face with
//                               rows padded to Nf' = a multiple of kFaceGroup
//                               faces; one monotonicity flag per face
//   tracer recon (K3 -> K4)     ReconRec (layer mean, interface values at the
//                               layer's top and bottom; FP64 in both forms)
//                               per layer, column and tracer of a group; each
//                               tracer's column contiguous by default
//   sublayer fluxes (K4 -> K5)  surface-major: packed targets [s][face] and
//                               fluxes [s][face][tracer], rows of Nf' faces
// The sweep reads each of its two columns once, front to back; its lanes
// are adjacent faces, which sit at similar depths, so with the column state
// interface-major per row the lanes read neighbouring records (and the
// column-state kernel, a lane per column, writes whole rows); its
// surface-major face-record writes are contiguous.  The face flux
// kernel gives each face one wavefront with a lane per neutral surface: a
// face's targets are non-decreasing in s, so the lanes walk down the two
// columns together and read consecutive per-column records.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "redi_column.hpp"
#include "redi_eos.hpp"

namespace redi {

struct alignas(16) ColRec {  // one interface of a column (K1 output)
  double T, S, dT, dS;        // interface T, S, dR/dT, dR/dS
};

// One neutral surface s of a face (K2 output), as the kernels use it:
// positions in the left / right column (in [0, 1]), effective thickness of
// sublayer s (m; 0 for s = ns-1), 1-based top-down layer indices.
struct SurfVals {
  double PoL, PoR, hEff;
  int KoL, KoR;
};

struct alignas(16) FaceRec64 {  // stored exactly
  double PoL, PoR, hEff;
  int32_t KoL, KoR;
  __device__ __forceinline__ SurfVals get() const { return SurfVals{PoL, PoR, hEff, KoL, KoR}; }
  static __device__ __forceinline__ FaceRec64 put(const SurfVals& v) {
    return FaceRec64{v.PoL, v.PoR, v.hEff, v.KoL, v.KoR};
  }
  static __device__ __forceinline__ double stored_heff(double h) { return h; }
};

// Positions as unsigned fractions of 2^31: 0 and 1 are exact and the
// absolute resolution is 2^-31 over the whole interval (an FP32 position
// near 1 resolves only 2^-24), in the same four bytes.  Sublayer
// interfaces sit at these positions, and the down-gradient guard compares
// interface values formed from them, so their resolution decides how
// often a guard of the FP32-store form decides differently from FP64.
struct alignas(16) FaceRec32 {
  uint32_t PoL, PoR;
  float hEff;
  int16_t KoL, KoR;
  static __device__ __forceinline__ uint32_t enc(double po) {
    return (uint32_t)__double2uint_rn(po * 0x1p31);
  }
  static __device__ __forceinline__ double dec(uint32_t q) { return (double)q * 0x1p-31; }
  __device__ __forceinline__ SurfVals get() const {
    return SurfVals{dec(PoL), dec(PoR), (double)hEff, KoL, KoR};
  }
  static __device__ __forceinline__ FaceRec32 put(const SurfVals& v) {
    return FaceRec32{enc(v.PoL), enc(v.PoR), (float)v.hEff, (int16_t)v.KoL, (int16_t)v.KoR};
  }
  static __device__ __forceinline__ double stored_heff(double h) { return (double)(float)h; }
};

// One layer of a tracer column, as the face kernel needs it: layer mean,
// limited top / bottom edge values, interface values at the layer's top /
// bottom.
struct LayerVals {
  double Tl, aL, aR, Tt, Tb;
};

// One layer of a tracer column (K3 output), FP64 in both forms: the layer
// mean, the PPM interface values at its top and bottom, and the branch the
// PPM limiter took.  Every branch yields the limited edges from exactly
// these three values (kept, flattened to the mean, or one edge rebuilt from
// the mean and the other), so the reader replays the branch on the same
// operands and obtains the same bits.  Every value a guard compares across
// a face is then exact in both forms -- the sweep places neutral surfaces
// where the T edges of the two columns nearly agree, so rounding them
// would flip guards.
struct alignas(16) ReconRec {
  double Tl, Tt, Tb;
  int32_t branch, pad;
  static __device__ __forceinline__ ReconRec make(double tl, double et, double eb) {
    double al = et, ar = eb;
    return ReconRec{tl, et, eb, ppm_limit<double>(tl, al, ar), 0};
  }
  __device__ __forceinline__ LayerVals get() const {
    double al, ar;
    ppm_limited_edges<double>(branch, Tl, Tt, Tb, al, ar);
    return LayerVals{Tl, al, ar, Tt, Tb};
  }
};

struct Fp64 {
  using R = double;  // stored fields and fluxes
  using Face = FaceRec64;
};
struct Fp32 {
  using R = float;
  using Face = FaceRec32;
};

template <class F> using FaceRecF = typename F::Face;

// Face record (face, s) lives at face*fa + s*fb: face-major, fa = ns,
// fb = 1.  A face kernel's wave reads its face's surfaces as one contiguous
// run; the sweep, whose lanes are adjacent faces, transposes its records
// through LDS to store whole runs.
inline uint32_t face_fa(int ns) { return (uint32_t)ns; }
inline uint32_t face_fb(int) { return 1u; }

// Flag bits per face: Ko of the contributing sublayers (hEff != 0) is
// non-decreasing in s on the left (bit 0) / right (bit 1) column.
constexpr uint8_t kMonoL = 1, kMonoR = 2;

// Thread -> cell/face mapping over TX x TY tiles of a 2-D index space, with
// tiles handed to the 8 XCDs in contiguous bands: workgroups are dispatched
// round-robin over XCDs, so workgroup b runs on XCD b % 8; giving each XCD a
// contiguous run of tiles keeps vertically adjacent tiles (which share
// v-faces and neighbour columns) in the same L2.
__device__ __forceinline__ int xcd_band(int bid, int nb) {
  const int q = nb >> 3, r = nb & 7, x = bid & 7, local = bid >> 3;
  return x * q + (x < r ? x : r) + local;
}

// Tiles are numbered along the fast axis first, or (vertical) along the
// slow axis first, so that an XCD's consecutive tiles are vertical
// neighbours: the v faces and columns two rows share are then read by
// consecutive tiles of the same L2.
constexpr int kTileVerticalDefault = 0;

struct TileMap {
  int n0, n1;    // index space extent (fast, slow)
  int tx, ty;    // tile shape (tx * ty == blockDim.x)
  int tiles0;    // tiles along the fast axis
  int tiles1;    // tiles along the slow axis
  int vertical;  // tile numbering: slow axis first
  // lanes of block `bid` of `nb` (tiles beyond the index space map outside it)
  __device__ __forceinline__ bool map(int bid, int nb, int& a, int& b) const {
    const int tile = xcd_band(bid, nb);
    const int t0 = vertical ? tile / tiles1 : tile % tiles0;
    const int t1 = vertical ? tile % tiles1 : tile / tiles0;
    a = t0 * tx + (int)(threadIdx.x % (unsigned)tx);
    b = t1 * ty + (int)(threadIdx.x / (unsigned)tx);
    return a < n0 && b < n1;
  }
};

// ---------------------------------------------------------------------------
// K1: column interface state (redi_build_column, redi/ocean_redi.F90:609-654)
// ---------------------------------------------------------------------------
// FP64 in both forms, the EOS included (the FP32-store form reads FP32
// fields): the sweep's density differences are differences of these
// derivatives' products, and FP32 derivatives move neutral surfaces into
// other layers.
// Column state layout: interface-major within each row of columns, the
// record of interface kt (top-down) of column (i, j) at (j*(nz+1) + kt)*nx + i.
// Neighbouring faces sit at similar depths, so the sweep's lanes (adjacent
// faces) read neighbouring records, and the column-state kernel's lanes
// (adjacent columns) write whole rows.
__device__ __forceinline__ size_t colstate_base(size_t c, int nx, int nz) {
  const size_t j = c / (size_t)nx, i = c % (size_t)nx;
  return j * (size_t)(nz + 1) * (size_t)nx + i;
}

template <class F>
struct ColumnArgs {
  int N2, nz, nx;
  const typename F::R* h;
  const typename F::R* thtr;
  const typename F::R* shtr;
  EosParams eos;
  ColRec* rec;
  double* P;
};

// Lane = column; `bid` is the workgroup's block index (blockDim.x lanes).
template <class F>
__device__ __forceinline__ void column_state_body(const ColumnArgs<F>& a, int bid) {
  const int c = bid * (int)blockDim.x + (int)threadIdx.x;
  if (c >= a.N2) return;
  const int nz = a.nz;
  const size_t n2 = (size_t)a.N2;
  const typename F::R* __restrict__ h = a.h;
  const typename F::R* __restrict__ thtr = a.thtr;
  const typename F::R* __restrict__ shtr = a.shtr;
  const EosParams eos = a.eos;
  const double g_rho0 = kGravity * eos.rho0;
  const size_t cb = colstate_base((size_t)c, a.nx, nz), cs = (size_t)a.nx;
  ColRec* r = a.rec + cb;
  double* pc = a.P + cb;
  double P = 0.0;  // Pint(1) = 0
  auto load = [&](int kt, double& hh, double* s) {
    const size_t o = (size_t)(nz - 1 - kt) * n2 + c;  // top-down -> native
    hh = (double)h[o];
    const double he = hh > kHDivEps ? hh : kHDivEps;
    s[0] = (double)thtr[o] / he;
    s[1] = (double)shtr[o] / he;
  };
  auto emit = [&](int k, const double* e, const double*) {
    if (k > 0) P = P + g_rho0 * (double)h[(size_t)(nz - k) * n2 + c];  // top-down layer k-1
    double dRdT, dRdS;
    interface_drho<double>(eos, e[0], e[1], P, dRdT, dRdS);
    r[(size_t)k * cs] = ColRec{e[0], e[1], dRdT, dRdS};
    pc[(size_t)k * cs] = P;
  };
  stream_edges<double, 2>(nz, load, emit);
}

template <class F>
__global__ void __launch_bounds__(128) k_column_state(ColumnArgs<F> a) {
  column_state_body<F>(a, (int)blockIdx.x);
}

// ---------------------------------------------------------------------------
// K2: continuous neutral-surface sweep
// (redi_neutral_positions_continuous, redi/ocean_redi.F90:307-464, run for
//  every wet face by redi_face_coeffs / redi_calc_coeffs_x/_y, :656-812)
// ---------------------------------------------------------------------------

// secant density difference rho(a) - rho(b) of two interface records
// (ocean_redi.F90:344-345, 367-371, 406-410)
__device__ __forceinline__ double sec_drho(const ColRec& a, const ColRec& b) {
  return 0.5 * ((a.dT + b.dT) * (a.T - b.T) + (a.dS + b.dS) * (a.S - b.S));
}

// Sliding window over one column of the sweep.  `k` is the reference's
// kl / kr (1-based, 1..nk+1); the window holds the records at m = max(k-1,1)
// and m+1 and the interface pressures at m, m+1, min(m+2, nk+1), which is
// every index the sweep reads while the pointer sits at k.
struct SweepSide {
  const ColRec* rec;
  const double* P;
  size_t st;  // elements between consecutive interfaces of the column
  int k, m;
  ColRec w0, w1;
  double p0, p1, p2;
  __device__ __forceinline__ void init(const ColRec* r, const double* p, size_t stride, int nk) {
    rec = r; P = p; st = stride; k = 1; m = 1;
    w0 = r[0]; w1 = r[stride];
    p0 = p[0]; p1 = p[stride]; p2 = p[(size_t)(nk < 2 ? nk : 2) * stride];
  }
  __device__ __forceinline__ ColRec cur() const { return k == 1 ? w0 : w1; }  // record at k
  __device__ __forceinline__ double pk() const { return k == 1 ? p0 : p1; }        // P(k)
  __device__ __forceinline__ double pk1() const { return k == 1 ? p1 : p2; }       // P(k+1)
  __device__ __forceinline__ void advance(int nk) {  // k -> k+1, k <= nk
    if (k >= 2) {
      m += 1;
      w0 = w1; w1 = rec[(size_t)m * st];
      p0 = p1; p1 = p2; p2 = P[(size_t)(m + 1 < nk ? m + 1 : nk) * st];
    }
    k += 1;
  }
};

template <class F>
struct SweepArgs {
  int nx, nz;
  int nfa;                 // faces per row of the face array
  uint32_t fa, fb;         // record (face, s) at face*fa + s*fb
  int i0, j0;              // first computed face
  int di, dj;              // left column = (i - di, j - dj)
  const double* wet;
  const ColRec* rec;
  const double* P;
  FaceRecF<F>* out;        // [s][face]
  uint8_t* mono;           // [face]
  double pa_to_h;          // 1 / (g rho0): Pa -> m (ocean_redi.F90:688-691)
  const double* inv_dx;    // 1 / face spacing per face (idxCu / idyCv): the slope taper's
  double slope_max;        // opt-in slope taper: 0 = off (the reference has none)
};

// Opt-in neutral-slope taper (no counterpart in the reference): sublayer
// between surfaces with absolute positions (pressures) top / bottom on the
// left and right columns has slope S = (zR - zL) / dx, z the mid-depth
// (mean of its bounding positions, Pa -> m); where |S| > slope_max its hEff
// is scaled by (slope_max / |S|)^2, the quadratic taper of Gerdes, Koberle
// and Willebrand (1991).  Applied to hEff in m before it is stored.
__device__ __forceinline__ double slope_taper(double hm, double topL, double botL, double topR,
                                              double botR, double pa_to_h, double inv_dx,
                                              double smax) {
  const double dz = ((0.5 * (topR + botR)) - (0.5 * (topL + botL))) * pa_to_h;
  const double aS = fabs(dz * inv_dx);
  if (aS > smax) {
    const double q = smax / aS;
    hm = hm * (q * q);
  }
  return hm;
}

// Density differences, positions, pressures and thicknesses are carried in
// FP64 in both forms; positions and thicknesses are stored in the form's
// type.
// The sweep's records reach the face-major face array through LDS: each
// lane puts its surface's record in its wave's stage, and every run of
// kSweepRunWords / (record words) surfaces the wave writes the stage out
// transposed -- 8 lanes per face, 8 faces per store instruction -- so every
// store fills whole 128-byte runs of records.
constexpr int kSweepRunWords = 8;                              // 16-byte words per face per run
constexpr int kSweepStageWords = 64 * kSweepRunWords;          // per wave
constexpr size_t kSweepStageBytes = kSweepStageWords * sizeof(uint4);

template <class F>
__device__ __forceinline__ void neutral_sweep_body(const SweepArgs<F>& a, const TileMap& tm,
                                                   int bid, int nb, uint4* stage) {
  using Face = FaceRecF<F>;
  constexpr int kW = (int)(sizeof(Face) / sizeof(uint4));
  constexpr int kRun = kSweepRunWords / kW;  // surfaces per run
  const int lane = (int)(threadIdx.x & 63);
  uint4* stw = stage + (threadIdx.x >> 6) * kSweepStageWords;
  int fi, fj;
  const bool inr = tm.map(bid, nb, fi, fj);
  const int i = a.i0 + fi, j = a.j0 + fj;
  const uint32_t fo = inr ? (uint32_t)i + (uint32_t)a.nfa * (uint32_t)j : 0u;
  const int nk = a.nz, ns = 2 * nk + 2;
  const bool wet = inr && a.wet[fo] > 0.0;  // dry face: the inert values of :722-730
  const SurfVals inert{0.0, 0.0, 0.0, 1, 1};
  auto put = [&](int s, const SurfVals& v) {
    const Face r = Face::put(v);
    uint4 w[kW];
    __builtin_memcpy(w, &r, sizeof(Face));
#pragma unroll
    for (int q = 0; q < kW; ++q) stw[lane * kSweepRunWords + (s % kRun) * kW + q] = w[q];
  };
  // surfaces s0 .. s0+cnt-1 of every face of the wave
  auto flush = [&](int s0, int cnt) {
    __syncthreads();
    const int wd = lane % kSweepRunWords, sq = wd / kW;
#pragma unroll
    for (int it = 0; it < 64 / (64 / kSweepRunWords); ++it) {
      const int q = it * (64 / kSweepRunWords) + lane / kSweepRunWords;
      const uint32_t foq = (uint32_t)__shfl((int)fo, q, 64);
      const int okq = __shfl(inr ? 1 : 0, q, 64);
      if (okq && sq < cnt) {
        uint4* dst = reinterpret_cast<uint4*>(a.out + (size_t)foq * a.fa + (size_t)(s0 + sq) * a.fb);
        dst[wd % kW] = stw[q * kSweepRunWords + wd];
      }
    }
    __syncthreads();
  };
  const size_t cl = (size_t)(i - a.di) + (size_t)a.nx * (j - a.dj);
  const size_t cr = (size_t)i + (size_t)a.nx * j;
  SweepSide L{}, Rt{};
  if (wet) {
    const size_t bl = colstate_base(cl, a.nx, nk), br = colstate_base(cr, a.nx, nk);
    L.init(a.rec + bl, a.P + bl, (size_t)a.nx, nk);
    Rt.init(a.rec + br, a.P + br, (size_t)a.nx, nk);
  }

  int lastK_l = 1, lastK_r = 1, contK_l = 0, contK_r = 0;
  double lastP_l = 0.0, lastP_r = 0.0, lastPos_l = 0.0, lastPos_r = 0.0;
  bool reached_bottom = false, search_l = false, mono_l = true, mono_r = true;
  SurfVals pend{};
  for (int ks = 1; ks <= ns; ++ks) {
   if (wet) {
    if (!reached_bottom) {
      const double dRho = sec_drho(Rt.cur(), L.cur());
      if (dRho < 0.0)
        search_l = true;
      else if (dRho > 0.0)
        search_l = false;
      else if (L.k + Rt.k == 2)
        search_l = true;
      else
        search_l = !search_l;
    }
    double PoL, PoR, posL, posR;
    int KoL, KoR;
    if (search_l) {
      const ColRec cr_ = Rt.cur();
      const double top = sec_drho(L.w0, cr_);
      const double bot = sec_drho(L.w1, cr_);
      double Po;
      if (top > 0.0 || Rt.k + L.k == 2)
        Po = 0.0;
      else if (top >= bot)
        Po = 1.0;
      else
        Po = interpolate_position<double>(top, L.p0, bot, L.p1);
      int Kn = L.m;
      double pa = L.p0, pb = L.p1;
      if (Po >= 1.0 && Kn < nk) {
        Kn += 1;
        Po = Po - 1.0;
        pa = L.p1;
        pb = L.p2;
      }
      if ((double)(Kn - lastK_l) + (Po - lastP_l) < 0.0) {  // monotonic backstop
        PoL = lastP_l; KoL = lastK_l; posL = lastPos_l;
      } else {
        PoL = Po; KoL = Kn; posL = pa + Po * (pb - pa);
      }
      if (Rt.k <= nk) {
        PoR = 0.0; KoR = Rt.k;
        const double q = Rt.pk();
        posR = q + 0.0 * (Rt.pk1() - q);
        Rt.advance(nk);
      } else {
        PoR = 1.0; KoR = nk;
        posR = Rt.p0 + 1.0 * (Rt.p1 - Rt.p0);
        reached_bottom = true; search_l = false;
      }
    } else {
      const ColRec cl_ = L.cur();
      const double top = sec_drho(Rt.w0, cl_);
      const double bot = sec_drho(Rt.w1, cl_);
      double Po;
      if (top >= 0.0 || Rt.k + L.k == 2)
        Po = 0.0;
      else if (top >= bot)
        Po = 1.0;
      else
        Po = interpolate_position<double>(top, Rt.p0, bot, Rt.p1);
      int Kn = Rt.m;
      double pa = Rt.p0, pb = Rt.p1;
      if (Po >= 1.0 && Kn < nk) {
        Kn += 1;
        Po = Po - 1.0;
        pa = Rt.p1;
        pb = Rt.p2;
      }
      if ((double)(Kn - lastK_r) + (Po - lastP_r) < 0.0) {
        PoR = lastP_r; KoR = lastK_r; posR = lastPos_r;
      } else {
        PoR = Po; KoR = Kn; posR = pa + Po * (pb - pa);
      }
      if (L.k <= nk) {
        PoL = 0.0; KoL = L.k;
        const double q = L.pk();
        posL = q + 0.0 * (L.pk1() - q);
        L.advance(nk);
      } else {
        PoL = 1.0; KoL = nk;
        posL = L.p0 + 1.0 * (L.p1 - L.p0);
        reached_bottom = true; search_l = true;
      }
    }
    lastK_l = KoL; lastP_l = PoL; lastK_r = KoR; lastP_r = PoR;
    if (ks > 1) {  // sublayer between surfaces ks-1 and ks (:452-462)
      const double hL = posL - lastPos_l;
      const double hR = posR - lastPos_r;
      double he = 0.0;
      if (hL + hR > 0.0) he = ((2.0 * hL) * hR) / (hL + hR);
      double hm = he * a.pa_to_h;
      if (a.slope_max > 0.0 && hm != 0.0)
        hm = slope_taper(hm, lastPos_l, posL, lastPos_r, posR, a.pa_to_h, a.inv_dx[fo], a.slope_max);
      pend.hEff = Face::stored_heff(hm);
      if (pend.hEff != 0.0) {
        mono_l = mono_l && pend.KoL >= contK_l;
        mono_r = mono_r && pend.KoR >= contK_r;
        contK_l = pend.KoL;
        contK_r = pend.KoR;
      }
      put(ks - 2, pend);
    }
    pend = SurfVals{PoL, PoR, 0.0, KoL, KoR};
    lastPos_l = posL;
    lastPos_r = posR;
   } else if (ks > 1) {
    put(ks - 2, inert);
   }
    if (ks > 1 && (ks - 2) % kRun == kRun - 1) flush(ks - 2 - (kRun - 1), kRun);
  }
  put(ns - 1, wet ? pend : inert);
  flush(ns - 1 - (ns - 1) % kRun, (ns - 1) % kRun + 1);
  if (inr) a.mono[fo] = wet ? (uint8_t)((mono_l ? kMonoL : 0) | (mono_r ? kMonoR : 0)) : (uint8_t)(kMonoL | kMonoR);
}

// Dynamic LDS: kSweepStageBytes per wave of the block.
template <class F>
__global__ void __launch_bounds__(256) k_neutral_sweep(SweepArgs<F> a, TileMap tm) {
  extern __shared__ uint4 sweep_stage[];
  neutral_sweep_body<F>(a, tm, (int)blockIdx.x, (int)gridDim.x, sweep_stage);
}

// ---------------------------------------------------------------------------
// K3: tracer column reconstruction (redi_tracer_column, :852-891), the NT
// tracers of a group per launch.  ReconRec (k, col, t) at col*ca + k*cb +
// t*ct (ReconLayout):
//   per column (ca = NT*nz, cb = 1, ct = nz): each tracer's column is one
//     run of records, which a face's lanes, walking down the column, read in
//     contiguous stretches;
//   layer-major (ca = NT, cb = N2*NT, ct = 1): the lanes of the streaming
//     kernel, adjacent columns, write contiguous rows.
// ---------------------------------------------------------------------------
template <class R>
struct TracerSetT {
  const R* q[4];  // thickness-weighted tracers of the group
};

// (32-bit element offsets from uniform bases: the host checks that
// (nz+1) * columns * 4 < 2^31.)
struct ReconLayout {
  uint32_t ca, cb, ct;
};

template <class F, int NT>
struct ReconArgs {
  int N2, nz;
  const typename F::R* h;
  TracerSetT<typename F::R> tr;
  ReconRec* rec;
  ReconLayout y;
};

// The reconstruction is computed in FP64 in both forms, from the stored
// fields.
// Streaming form: one lane per column walks it top-down with a register
// window (stream_edges).
template <class F, int NT>
__global__ void __launch_bounds__(128) k_tracer_recon(ReconArgs<F, NT> a) {
  const int N2 = a.N2, nz = a.nz;
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= N2) return;
  const size_t n2 = (size_t)N2;
  const typename F::R* __restrict__ h = a.h;
  auto load = [&](int kt, double& hh, double* s) {
    const size_t o = (size_t)(nz - 1 - kt) * n2 + c;
    hh = (double)h[o];
    const double he = hh > kHDivEps ? hh : kHDivEps;
#pragma unroll
    for (int t = 0; t < NT; ++t) s[t] = (double)a.tr.q[t][o] / he;
  };
  double e_prev[NT];
  ReconRec* rc = a.rec + (size_t)c * a.y.ca;
  auto emit = [&](int k, const double* e, const double* s_above) {
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      if (k > 0) rc[(size_t)(k - 1) * a.y.cb + (size_t)t * a.y.ct] = ReconRec::make(s_above[t], e_prev[t], e[t]);
      e_prev[t] = e[t];
    }
  };
  stream_edges<double, NT>(nz, load, emit);
}

// Tiled form: a workgroup of kReconLanes lanes takes CT = kReconLanes / KT
// adjacent columns x KT layers (KT, the layer chunk, is 8, 16 or 32).  Layer
// values, PLM slopes and PPM edges are staged in LDS stage by stage (every
// value from the same functions on the same operands as stream_edges,
// including its clamped windows at the top and bottom), so global reads run
// along columns and the per-column record runs are written along layers.
// The chunk changes only the schedule, never a result.  Per-column layout
// only (cb = 1).
constexpr int kReconLanes = 256;
constexpr int kReconChunkDefault = 16;


// LDS doubles of one tiled-reconstruction workgroup
template <int NT, int KT>
constexpr int recon_lds_doubles() {
  constexpr int CT = kReconLanes / KT;
  return (KT + 4) * CT * (1 + NT) + NT * (KT + 2) * CT + NT * (KT + 1) * CT;
}
template <int NT>
constexpr int recon_lds_max() {
  constexpr int a = recon_lds_doubles<NT, 8>(), b = recon_lds_doubles<NT, 16>(),
                c = recon_lds_doubles<NT, 32>();
  return a > b ? (a > c ? a : c) : (b > c ? b : c);
}

template <class F, int NT, int KT>
__device__ __forceinline__ void tracer_recon_tiled_body(const ReconArgs<F, NT>& a, int bid, int nb,
                                                        double* lds) {
  using R = double;
  constexpr int CT = kReconLanes / KT, NL = KT + 4, NSL = KT + 2, NE = KT + 1;
  // h, layers k0-2 .. k0+KT+1 (clamped); layer values hTr / max(h, eps), same
  // layers; PLM slopes, layers k0-1 .. k0+KT; PPM edges, interfaces k0 .. k0+KT
  R(*sh)[CT] = reinterpret_cast<R(*)[CT]>(lds);
  R(*sv)[NL][CT] = reinterpret_cast<R(*)[NL][CT]>(lds + NL * CT);
  R(*sd)[NSL][CT] = reinterpret_cast<R(*)[NSL][CT]>(lds + NL * CT * (1 + NT));
  R(*se)[NE][CT] = reinterpret_cast<R(*)[NE][CT]>(lds + NL * CT * (1 + NT) + NT * NSL * CT);
  const int N2 = a.N2, nz = a.nz;
  const typename F::R* __restrict__ h = a.h;
  const int nkb = (nz + KT - 1) / KT;
  const int b = xcd_band(bid, nb);
  const int c0 = (b / nkb) * CT, k0 = (b % nkb) * KT;
  const int tid = (int)threadIdx.x;
  const size_t n2 = (size_t)N2;
  // 1: layer values, lanes along columns
  for (int q = tid; q < NL * CT; q += kReconLanes) {
    const int r = q / CT, cc = q % CT;
    int kl = k0 - 2 + r;
    kl = kl < 0 ? 0 : (kl > nz - 1 ? nz - 1 : kl);
    const int c = c0 + cc < N2 ? c0 + cc : N2 - 1;
    const size_t o = (size_t)(nz - 1 - kl) * n2 + c;  // top-down layer kl
    const R hh = (R)h[o];
    const R he = hh > R(kHDivEps) ? hh : R(kHDivEps);
    sh[r][cc] = hh;
#pragma unroll
    for (int t = 0; t < NT; ++t) sv[t][r][cc] = (R)a.tr.q[t][o] / he;
  }
  __syncthreads();
  // 2: slopes of layers kl = k0-1+r; 0 for the first and last layer (the
  // slope stream_edges starts from and the one it skips)
  for (int q = tid; q < NSL * CT; q += kReconLanes) {
    const int r = q / CT, cc = q % CT;
    const int kl = k0 - 1 + r;  // rows r, r+1, r+2 of sh / sv hold layers kl-1, kl, kl+1
    const bool in = kl >= 1 && kl <= nz - 2;
#pragma unroll
    for (int t = 0; t < NT; ++t)
      sd[t][r][cc] = in ? plm_slope<R>(sh[r][cc], sh[r + 1][cc], sh[r + 2][cc], sv[t][r][cc],
                                       sv[t][r + 1][cc], sv[t][r + 2][cc])
                        : R(0);
  }
  __syncthreads();
  // 3: edges at interfaces ki = k0+r
  for (int q = tid; q < NE * CT; q += kReconLanes) {
    const int r = q / CT, cc = q % CT;
    const int ki = k0 + r;  // sh / sv rows r..r+3: layers ki-2 .. ki+1 (clamped);
                            // sd rows r, r+1: slopes of layers ki-1, ki
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      R e = R(0);
      if (ki == 0) {
        e = sv[t][r + 2][cc] - R(0.5) * R(0);  // edge(1) = tr(1) - 0.5*diff(1), diff(1) = 0
      } else if (ki == nz) {
        e = sv[t][r + 1][cc] + R(0.5) * R(0);  // diff(nk) = 0
      } else if (ki < nz) {
        e = ppm_edge<R>(sh[r][cc], sh[r + 1][cc], sh[r + 2][cc], sh[r + 3][cc], sv[t][r + 1][cc],
                        sv[t][r + 2][cc], sd[t][r][cc], sd[t][r + 1][cc]);
      }
      se[t][r][cc] = e;
    }
  }
  __syncthreads();
  // 4: records of layers kl = k0+kk, lanes along layers
  const int kk = tid % KT, cc = tid / KT;
  const int kl = k0 + kk, c = c0 + cc;
  if (kl < nz && c < N2) {
    ReconRec* rc = a.rec + (size_t)c * a.y.ca + (size_t)kl * a.y.cb;
#pragma unroll
    for (int t = 0; t < NT; ++t)
      rc[(size_t)t * a.y.ct] = ReconRec::make(sv[t][kk + 2][cc], se[t][kk][cc], se[t][kk + 1][cc]);
  }
}

template <class F, int NT, int KT>
__global__ void __launch_bounds__(kReconLanes) k_tracer_recon_tiled(ReconArgs<F, NT> a) {
  __shared__ double lds[recon_lds_doubles<NT, KT>()];
  tracer_recon_tiled_body<F, NT, KT>(a, (int)blockIdx.x, (int)gridDim.x, lds);
}

// Workgroups of the tiled reconstruction of one domain.
inline unsigned recon_tiled_blocks(int N2, int nz, int kt) {
  const int ct = kReconLanes / kt;
  return (unsigned)((N2 + ct - 1) / ct) * (unsigned)((nz + kt - 1) / kt);
}

// ---------------------------------------------------------------------------
// K4: sublayer fluxes, once per face (redi_face_flux + redi_sublayer_dT,
// :893-925, :1037-1092)
// ---------------------------------------------------------------------------

// Pieces of redi_sublayer_dT (:893-925).  The interface value at position
// Po in layer K of a column, (1 - Po) * T(K-1) + Po * T(K), from layer K's
// record (the tlt / tlb / trt / trb expressions, :905-912):
__device__ __forceinline__ double edge_value(const LayerVals& r, double Po) {
  return (1.0 - Po) * r.Tt + Po * r.Tb;
}

// The down-gradient guard (:913-924): the layer difference where the
// top, bottom, mean and layer differences agree in sign, else 0.
template <class R>
__device__ __forceinline__ R guarded_dT(R tlt, R tlb, R trt, R trb, R tlay, R trlay) {
  const R dT_top = trt - tlt;
  const R dT_bot = trb - tlb;
  const R dT_ave = R(0.5) * (dT_top + dT_bot);
  const R dT_layer = trlay - tlay;
  if (signum1<R>(dT_top) * signum1<R>(dT_bot) <= R(0) ||
      signum1<R>(dT_ave) * signum1<R>(dT_layer) <= R(0))
    return R(0);
  return dT_layer;
}

constexpr int kMaxNT = 4;

// A face's sublayer ks contributes when hEff(ks) != 0, to layer KoL(ks) of
// its left column and KoR(ks) of its right column (1-based, top-down, at
// most 256).  Its targets are stored packed as KoL | KoR << 16; 0 marks a
// sublayer that contributes nothing.
constexpr int kTargetShiftL = 0, kTargetShiftR = 16;

// The fluxes of one face grid (u or v).  One wavefront per face, a
// workgroup takes kFaceGroup consecutive faces.
constexpr int kFaceGroup = 4;
constexpr int kChunk = 63;  // sublayers per pass of a wavefront (64 surfaces)
// A face wave gathers the reconstruction records of kFaceSub tracers at a
// time, which fits the kernel in 80 VGPRs and kFaceWaves waves per SIMD.
// The kernel's cost is mostly the latency of its dependent loads, so more
// waves beat more loads in flight per wave.  Measured on MI355X at four
// tracers: holding all four tracers' records, 126 VGPRs, four waves; two at
// a time, 96, five waves; one at a time, six waves, 17% (FP64) and 12%
// (FP32-store) faster than four; seven waves spill.
constexpr int kFaceSub = 1, kFaceWaves = 6;

template <class F>
struct FaceFluxArgs {
  int nfa;                // faces per row of the face array
  int ilo, ihi, jlo, jhi; // faces computed: i in [ilo, ihi), j in [jlo, jhi) (0-based)
  int di, dj;             // left column = (i - di, j - dj), right column = (i, j)
  int nx, ns;             // columns per row, neutral surfaces
  int nzc;                // layers per column (ns = 2*nzc + 2)
  int wall0, wall1;       // 0-based face coordinate across the faces (i for u,
                          // j for v) of the two wall faces, -1 where open
  uint32_t nf;            // faces in the face array
  uint32_t fa, fb;        // face record (face, s) at face*fa + s*fb
  uint32_t nfp;           // faces per sublayer row of tgt / flx (nf rounded up to kFaceGroup)
  double dt;
  const double *khtr, *len, *inv;  // face diffusivity, face length, 1 / spacing
  const FaceRecF<F>* rec;
  ReconLayout ry;
  const ReconRec* recon;      // tracer t's records from recon + t*ry.ct
  uint32_t* tgt;              // packed targets, (s, face) at s*nfp + face
  typename F::R* flx;         // fluxes, (s, face, t) at (s*nfp + face)*NT + t
};

// flx = dT * hEff * Coef with Coef = dt*khtr*dy_cu*idxCu (v: dt*khtr*dx_cv*idyCv)
// for every contributing sublayer, in the reference's rounding order
// (:1060-1090), computed in FP64 in both forms and stored in the form's type.  The reference evaluates this twice per face, once from
// each bounding cell; both evaluations are the same expression on the same
// operands, so one evaluation serves both cells.
// Lane l of a face's wavefront takes neutral surface s = c0 + l: it loads
// the surface's record and, on each side, the record of the layer the
// surface lies in, and forms the interface value there.  Sublayer s lies
// between surfaces s and s+1, so lane l (< 63) takes the bottom-surface
// values from lane l+1 by a wavefront shuffle; they are the same
// expressions the reference evaluates for the sublayer's bottom.  A face's
// targets are non-decreasing in s, so the lanes read consecutive records of
// each column.  The kFaceGroup waves of a workgroup read one line of each
// surface-major face-record row between them; their fluxes are staged in
// LDS and written back as whole rows (kFaceGroup faces x NT tracers).
// sx: kFaceGroup * 64 * NT fluxes (stored type), stg: kFaceGroup * 64
// packed targets, both in LDS.
// Staged record reads (kStage, per-column reconstruction order only): a
// chunk's 64 surfaces lie in few layers of each column -- about half as many
// as surfaces -- so instead of every lane gathering its own two records per
// tracer, the wavefront loads the layer range [lo, hi] each side touches
// with consecutive lanes (one coalesced pass per side and tracer) into 64
// LDS records per side, and each lane reads its records there.  A chunk
// whose range on either side spans more than 64 layers takes the gather.
// The records read are the same, so are the bits.
constexpr int kStageRecs = 128;  // LDS records per wavefront (64 per side)
constexpr bool kFaceStageDefault = false;  // staged reads unless a caller chooses

// Face group `grp` (faces grp*kFaceGroup .. +kFaceGroup-1 of the face array).
template <class F, int NT, bool kStage = false>
__device__ __forceinline__ void face_flux_group(const FaceFluxArgs<F>& a, int grp,
                                                typename F::R* sx, uint32_t* stg,
                                                ReconRec* stage_lds = nullptr) {
  using R = typename F::R;
  using Face = FaceRecF<F>;
  const int wv = (int)(threadIdx.x >> 6), lane = (int)(threadIdx.x & 63);
  const size_t f0 = (size_t)grp * kFaceGroup;
  const size_t fo = f0 + wv;
  const int i = (int)(fo % (size_t)a.nfa), j = (int)(fo / (size_t)a.nfa);
  const int across = a.di ? i : j;
  const bool live = fo < a.nf && i >= a.ilo && i < a.ihi && j >= a.jlo && j < a.jhi &&
                    across != a.wall0 && across != a.wall1;
  uint32_t cl = 0, cr = 0;
  double coef = 0.0;
  if (live) {
    cr = (uint32_t)i + (uint32_t)a.nx * (uint32_t)j;
    cl = cr - (uint32_t)a.di - (uint32_t)a.nx * (uint32_t)a.dj;
    coef = ((a.dt * a.khtr[fo]) * a.len[fo]) * a.inv[fo];
  }
  const ReconRec* recL = a.recon + cl * a.ry.ca;
  const ReconRec* recR = a.recon + cr * a.ry.ca;
  const uint32_t cb = a.ry.cb, ct = a.ry.ct;
  const size_t fb = a.fb;
  const int ns = a.ns, nsub = ns - 1;
  const Face* r = a.rec + (live ? fo * a.fa : 0);
  for (int c0 = 0; c0 < nsub; c0 += kChunk) {
    const int s = c0 + lane;
    const bool surf = live && s < ns;
    SurfVals Fs{0.0, 0.0, 0.0, 1, 1};
    if (surf) Fs = r[(size_t)s * fb].get();
    const int KoL = Fs.KoL, KoR = Fs.KoR;
    const double PoL = Fs.PoL, PoR = Fs.PoR;
    bool staged = false;
    int lo_l = 0, hi_l = 0, lo_r = 0, hi_r = 0;
    if constexpr (kStage) {
      // wave-uniform layer ranges of this chunk's surfaces: positions Ko + Po
      // never decrease along the surfaces (the sweep's backstop), so every
      // surface's layer lies within one of the first and last surfaces';
      // a ballot confirms it, else the chunk gathers
      const int last = (live ? (ns - c0 < 64 ? ns - c0 : 64) : 1) - 1;
      const int kl0 = __builtin_amdgcn_readlane(KoL, 0), kl1 = __builtin_amdgcn_readlane(KoL, last);
      const int kr0 = __builtin_amdgcn_readlane(KoR, 0), kr1 = __builtin_amdgcn_readlane(KoR, last);
      lo_l = kl0 > 1 ? kl0 - 1 : 1;
      hi_l = kl1 + 1;
      lo_r = kr0 > 1 ? kr0 - 1 : 1;
      hi_r = kr1 + 1;
      const bool out = surf && (KoL < lo_l || KoL > hi_l || KoR < lo_r || KoR > hi_r);
      staged = live && hi_l - lo_l < 64 && hi_r - lo_r < 64 && __ballot(out) == 0;
      if (hi_l > a.nzc) hi_l = a.nzc;  // stage no layer past the column
      if (hi_r > a.nzc) hi_r = a.nzc;
    }
    // surface s+1, the bottom of sublayer s
    const double PoLb = __shfl_down(PoL, 1, 64), PoRb = __shfl_down(PoR, 1, 64);
    const int KoLb = __shfl_down(KoL, 1, 64), KoRb = __shfl_down(KoR, 1, 64);
    const bool contrib = surf && lane < kChunk && s < nsub && Fs.hEff != 0.0;
    const uint32_t packed =
        contrib ? (((uint32_t)KoL << kTargetShiftL) | ((uint32_t)KoR << kTargetShiftR)) : 0u;
    const double hEff = Fs.hEff;
    R x[NT];
    // tracers in sub-groups of kFaceSub: a sub-group's records are gathered
    // and its fluxes formed before the next sub-group's records are loaded,
    // so a wave holds the records of kFaceSub tracers at a time
#pragma unroll
    for (int g0 = 0; g0 < NT; g0 += kFaceSub) {
      LayerVals qL[kFaceSub], qR[kFaceSub];
      double eL[kFaceSub], eR[kFaceSub];
#pragma unroll
      for (int u = 0; u < kFaceSub; ++u) {
        const int t = g0 + u;
        if (t >= NT) break;
        qL[u] = LayerVals{0.0, 0.0, 0.0, 0.0, 0.0};
        qR[u] = qL[u];
        if (kStage && staged) {
          ReconRec* sl = stage_lds + wv * kStageRecs;
          if (lane <= hi_l - lo_l) sl[lane] = recL[(uint32_t)(lo_l - 1 + lane) + (uint32_t)t * ct];
          if (lane <= hi_r - lo_r) sl[64 + lane] = recR[(uint32_t)(lo_r - 1 + lane) + (uint32_t)t * ct];
          // a wavefront's LDS accesses execute in issue order, so only the
          // compiler must keep the reads after the writes (and the next
          // tracer's writes after these reads)
          __builtin_amdgcn_wave_barrier();
          asm volatile("" ::: "memory");
          if (surf) {
            qL[u] = sl[KoL - lo_l].get();
            qR[u] = sl[64 + KoR - lo_r].get();
          }
          asm volatile("" ::: "memory");
          __builtin_amdgcn_wave_barrier();
        } else if (surf) {
          qL[u] = recL[(uint32_t)(KoL - 1) * cb + (uint32_t)t * ct].get();
          qR[u] = recR[(uint32_t)(KoR - 1) * cb + (uint32_t)t * ct].get();
        }
        eL[u] = edge_value(qL[u], PoL);
        eR[u] = edge_value(qR[u], PoR);
      }
      double eLb[kFaceSub], eRb[kFaceSub];
#pragma unroll
      for (int u = 0; u < kFaceSub; ++u) {
        if (g0 + u >= NT) break;
        eLb[u] = __shfl_down(eL[u], 1, 64);
        eRb[u] = __shfl_down(eR[u], 1, 64);
      }
#pragma unroll
      for (int u = 0; u < kFaceSub; ++u) {
        const int t = g0 + u;
        if (t >= NT) break;
        x[t] = R(0);
        if (contrib) {
          const double tlay = ppm_ave<double>(PoL, PoLb + (double)(KoLb - KoL), qL[u].aL, qL[u].aR, qL[u].Tl);
          const double trlay = ppm_ave<double>(PoR, PoRb + (double)(KoRb - KoR), qR[u].aL, qR[u].aR, qR[u].Tl);
          const double d = guarded_dT<double>(eL[u], eLb[u], eR[u], eRb[u], tlay, trlay);
          x[t] = (R)((d * hEff) * coef);
        }
      }
      // the next sub-group's loads stay after this one's fluxes
      if (g0 + kFaceSub < NT) asm volatile("" ::: "memory");
    }
    // the chunk's rows in LDS in their global order: row rr holds the
    // group's faces q, each with its NT fluxes
#pragma unroll
    for (int t = 0; t < NT; ++t) sx[(lane * kFaceGroup + wv) * NT + t] = x[t];
    stg[lane * kFaceGroup + wv] = packed;
    __syncthreads();
    // each row (kFaceGroup faces x NT fluxes, a multiple of 16 bytes) is one
    // contiguous stretch of the surface-major flux array: the workgroup
    // writes the rows as 16-byte words, consecutive threads taking
    // consecutive words, and each row's packed targets as one word
    const int rows = nsub - c0 < kChunk ? nsub - c0 : kChunk;
    constexpr int kRowWords = (int)(kFaceGroup * NT * sizeof(R) / sizeof(uint4));
    const uint4* sxw = reinterpret_cast<const uint4*>(sx);
    for (int k = (int)threadIdx.x; k < rows * kRowWords; k += 64 * kFaceGroup) {
      const int row = k / kRowWords, w = k % kRowWords;
      uint4* dst = reinterpret_cast<uint4*>(a.flx + ((size_t)(c0 + row) * a.nfp + f0) * NT);
      dst[w] = sxw[k];
    }
    if ((int)threadIdx.x < rows)
      *reinterpret_cast<uint4*>(a.tgt + (size_t)(c0 + threadIdx.x) * a.nfp + f0) =
          reinterpret_cast<const uint4*>(stg)[threadIdx.x];
    __syncthreads();
  }
}

// Block bid of nb over one face grid: consecutive groups per XCD.
template <class F, int NT, bool kStage = false>
__device__ __forceinline__ void face_flux_body(const FaceFluxArgs<F>& a, int bid, int nb,
                                               typename F::R* sx, uint32_t* stg,
                                               ReconRec* stage_lds = nullptr) {
  face_flux_group<F, NT, kStage>(a, xcd_band(bid, nb), sx, stg, stage_lds);
}

template <class F, int NT>
__global__ void __launch_bounds__(64 * kFaceGroup, kFaceWaves) k_face_flux(FaceFluxArgs<F> a) {
  __shared__ alignas(16) typename F::R sx[kFaceGroup * 64 * NT];
  __shared__ alignas(16) uint32_t stg[kFaceGroup * 64];
  face_flux_body<F, NT>(a, (int)blockIdx.x, (int)gridDim.x, sx, stg);
}

// The same with staged record reads (per-column reconstruction order).
template <class F, int NT>
__global__ void __launch_bounds__(64 * kFaceGroup) k_face_flux_staged(FaceFluxArgs<F> a) {
  __shared__ alignas(16) typename F::R sx[kFaceGroup * 64 * NT];
  __shared__ alignas(16) uint32_t stg[kFaceGroup * 64];
  __shared__ ReconRec stage[kFaceGroup * kStageRecs];
  face_flux_body<F, NT, true>(a, (int)blockIdx.x, (int)gridDim.x, sx, stg, stage);
}

// ---------------------------------------------------------------------------
// K5: flux divergence per cell (redi_apply_flux_impl, :1094-1181)
// ---------------------------------------------------------------------------
template <class F>
struct CellFluxArgs {
  int nx, ny, nz;
  int wuf_w, wuf_e, wvf_s, wvf_n;  // 1-based wall faces (:1139-1142)
  int wall_w, wall_e, wall_s, wall_n;
  const double* areaT;
  const uint32_t *tgt_u, *tgt_v;
  const typename F::R *flx_u, *flx_v;
  uint32_t nu, nv;                 // faces per sublayer row of the u / v arrays
  int vertical;                    // segments numbered down the rows first
  typename F::R* htr[kMaxNT];
  double* tend[kMaxNT];            // optional: the FP64 tendency dTr / areaT per cell
};

// A pointer the compiler cannot trace to a kernel argument (the batched
// kernels read theirs from a work table) is dereferenced with flat
// instructions, which also count against the LDS wait counter: every LDS
// wait then waits for all loads in flight.  Every device array here is in
// global memory; pointers of this type keep the accesses global.
template <class T>
using Global = __attribute__((address_space(1))) T;
template <class T>
__device__ __forceinline__ Global<T>* global_ptr(T* p) {
  return (Global<T>*)p;
}

// Layers an accumulator column of k_cell_flux holds.
constexpr int kCellKCap[4] = {32, 76, 128, 256};

// One wavefront per workgroup; lane = (cell, tracer), 64 / NT consecutive
// cells of one row.  dTr(1:nz) of each lane lives in LDS in FP64 (both
// forms), one column per lane (acc[K][lane]), so a lane can add into any
// layer: every open face of the cell (W, E, S, N) is streamed sublayer by
// sublayer (s ascending) and each contributing sublayer's flux is added
// into dTr(nz+1-Ko(s)), exactly the order of the reference's loops,
// whatever the pattern of Ko.  The lanes of a wavefront read one contiguous
// stretch of each surface-major sublayer row.  Then hTr += dTr * (1/areaT).
// acc: KCAP * 64 doubles of LDS.
template <class F, int NT, int KCAP>
__device__ __forceinline__ void cell_flux_body(const CellFluxArgs<F>& a, int bid, int nb,
                                               double* acc) {
  using R = typename F::R;
  constexpr int kCells = 64 / NT;  // cells per wavefront
  // sublayer rows per batch: columns of at most 32 layers (65 sublayers)
  // take a face in two or three batches of 32, deeper ones stream batches
  // of 16
  constexpr int kBatch = KCAP <= 32 ? 32 : 16;
  const int lane = (int)threadIdx.x;
  const int cw = lane / NT, t = lane % NT;
  const int segs = (a.nx + kCells - 1) / kCells;
  const int w = xcd_band(bid, nb);
  const int cj = a.vertical ? w % a.ny : w / segs;
  const int ci = (a.vertical ? w / a.ny : w % segs) * kCells + cw;
  const bool cell = cw < kCells && ci < a.nx && cj < a.ny;
  const int nz = a.nz, nsub = 2 * nz + 1;
  for (int K = 0; K < nz; ++K) acc[K * 64 + lane] = 0.0;
  if (!cell) return;  // lanes only ever touch their own accumulator column
  const int nx = a.nx, ny = a.ny;
  const int i = ci + 1, j = cj + 1;  // 1-based
  // the four faces: W (u face i, this cell on the right), E (u face i+1,
  // left), S (v face j, right), N (v face j+1, left) (:1136-1175)
  const bool act[4] = {
      i >= 2 && !((a.wall_w && i == a.wuf_w) || (a.wall_e && i == a.wuf_e)),
      i <= nx - 1 && !((a.wall_w && i + 1 == a.wuf_w) || (a.wall_e && i + 1 == a.wuf_e)),
      j >= 2 && !((a.wall_s && j == a.wvf_s) || (a.wall_n && j == a.wvf_n)),
      j <= ny - 1 && !((a.wall_s && j + 1 == a.wvf_s) || (a.wall_n && j + 1 == a.wvf_n))};
  const size_t fo[4] = {(size_t)(i - 1) + (size_t)(nx + 1) * (j - 1),
                        (size_t)i + (size_t)(nx + 1) * (j - 1),
                        (size_t)(i - 1) + (size_t)nx * (j - 1),
                        (size_t)(i - 1) + (size_t)nx * j};
#pragma unroll 1
  for (int f = 0; f < 4; ++f) {
    if (!act[f]) continue;
    const bool left = (f & 1) != 0;
    const int shift = left ? kTargetShiftL : kTargetShiftR;
    const size_t nf = f < 2 ? a.nu : a.nv;
    const Global<const uint32_t>* tg = global_ptr(f < 2 ? a.tgt_u : a.tgt_v) + fo[f];
    const Global<const R>* fx = global_ptr(f < 2 ? a.flx_u : a.flx_v) + fo[f] * NT + t;
    // dTr(K) of the run of consecutive contributions to one layer K stays
    // in a register; it is read from / written back to LDS when K changes,
    // so every addition still lands in reference order
    int curK = 0;
    double cur = 0.0;
    // every (sublayer, face) slot of a computed face holds a flux (0 where
    // nothing contributes), so both loads issue without waiting; rows past
    // the last sublayer re-read the last one and are skipped, which keeps
    // the loads free of branches
    auto load = [&](int s0, uint32_t (&pk)[kBatch], R (&x)[kBatch]) {
#pragma unroll
      for (int u = 0; u < kBatch; ++u) {
        const int s = s0 + u < nsub ? s0 + u : nsub - 1;
        pk[u] = tg[(size_t)s * nf];
        x[u] = fx[(size_t)s * nf * NT];
      }
    };
    auto add = [&](int s0, const uint32_t (&pk)[kBatch], const R (&x)[kBatch]) {
#pragma unroll
      for (int u = 0; u < kBatch; ++u) {
        if (s0 + u >= nsub || pk[u] == 0) continue;
        const int K = (int)((pk[u] >> shift) & 0xffffu);
        if (K != curK) {
          if (curK) acc[(curK - 1) * 64 + lane] = cur;
          cur = acc[(K - 1) * 64 + lane];
          curK = K;
        }
        cur = left ? cur + (double)x[u] : cur - (double)x[u];
      }
    };
    // two batches of rows in registers: the next batch's loads are in
    // flight while the current one is added (one wave per SIMD, so nothing
    // else hides the latency).  Two batches of 16 are 64 loads, about the
    // 63 a wave's load counter tracks.  In the steady loop the loads are issued
    // unconditionally (no branch around them), so that the wait for one
    // batch never has to wait for the batch behind it (measured on MI355X:
    // 10.8 against 11.8 ms with conditional loads of 32-row batches, FP64);
    // the last batches, which may be past the last sublayer, come after it
    uint32_t pa[kBatch], pb[kBatch];
    R xa[kBatch], xb[kBatch];
    load(0, pa, xa);
    if constexpr (KCAP <= 32) {  // at most three batches: no steady loop
      for (int s0 = 0; s0 < nsub; s0 += 2 * kBatch) {
        const bool second = s0 + kBatch < nsub;
        if (second) load(s0 + kBatch, pb, xb);
        add(s0, pa, xa);
        if (s0 + 2 * kBatch < nsub) load(s0 + 2 * kBatch, pa, xa);
        if (second) add(s0 + kBatch, pb, xb);
      }
    } else {
      int s0 = 0;
      for (; s0 + 2 * kBatch < nsub; s0 += 2 * kBatch) {
        load(s0 + kBatch, pb, xb);
        add(s0, pa, xa);
        load(s0 + 2 * kBatch, pa, xa);
        add(s0 + kBatch, pb, xb);
      }
      if (s0 + kBatch < nsub) {  // the last one or two batches
        load(s0 + kBatch, pb, xb);
        add(s0, pa, xa);
        add(s0 + kBatch, pb, xb);
      } else {
        add(s0, pa, xa);
      }
    }
    if (curK) acc[(curK - 1) * 64 + lane] = cur;
  }
  const size_t n2 = (size_t)nx * ny;
  const size_t c = (size_t)ci + (size_t)nx * cj;
  const double iaij = 1.0 / global_ptr(a.areaT)[c];
  R* hq = a.htr[0];
  double* tq = a.tend[0];
#pragma unroll
  for (int q = 1; q < NT; ++q)
    if (t == q) { hq = a.htr[q]; tq = a.tend[q]; }
  Global<R>* hp = global_ptr(hq);
  Global<double>* tp = global_ptr(tq);
  for (int K = 1; K <= nz; ++K) {
    const size_t o = (size_t)(nz - K) * n2 + c;  // native layer nz+1-K (1-based)
    const double d = acc[(K - 1) * 64 + lane] * iaij;
    hp[o] = (R)((double)hp[o] + d);
    if (tp) tp[o] = d;
  }
}

template <class F, int NT, int KCAP>
__global__ void __launch_bounds__(64) k_cell_flux(CellFluxArgs<F> a) {
  __shared__ double acc[KCAP * 64];
  cell_flux_body<F, NT, KCAP>(a, (int)blockIdx.x, (int)gridDim.x, acc);
}

// Workgroups of k_cell_flux for an nx x ny domain.
template <int NT>
inline unsigned cell_flux_blocks(int nx, int ny) {
  constexpr int kCells = 64 / NT;
  return (unsigned)ny * (unsigned)((nx + kCells - 1) / kCells);
}

// ---------------------------------------------------------------------------
// Face records <-> reference arrays, khtr faces, initial fill
// ---------------------------------------------------------------------------
struct CoeffArrays {  // the reference layout: [s][face] (Fortran (nfa,nfb,ns))
  double *PoL, *PoR, *hEff;
  int32_t *KoL, *KoR;
};

// Faces (i, j), i in [i0, i1), j in [j0, j1) of an nfa-wide face array.
template <class F>
__global__ void k_export_coeffs(const FaceRecF<F>* __restrict__ rec, uint32_t fa, uint32_t fb,
                                int nfa, int nfb, int ns, int i0, int i1, int j0, int j1,
                                CoeffArrays out) {
  const size_t nf = (size_t)nfa * nfb;
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nf * ns) return;
  const size_t f = t % nf, s = t / nf;  // writes coalesced along the face index
  const int i = (int)(f % nfa), j = (int)(f / nfa);
  if (i < i0 || i >= i1 || j < j0 || j >= j1) return;
  const SurfVals r = rec[f * fa + s * fb].get();
  out.PoL[t] = r.PoL;
  out.PoR[t] = r.PoR;
  out.KoL[t] = r.KoL;
  out.KoR[t] = r.KoR;
  if ((int)s < ns - 1) out.hEff[t] = r.hEff;
}

struct KhtrArgs {
  int nu, nv;
  const double* ext_u;  // NULL: the scalar khtr
  const double* ext_v;
  double khtr;
  double* ku;
  double* kv;
};

__device__ __forceinline__ void face_khtr_body(const KhtrArgs& a, int bid) {
  const int t = bid * (int)blockDim.x + (int)threadIdx.x;
  if (t < a.nu) {
    a.ku[t] = a.ext_u ? a.ext_u[t] : a.khtr;  // redi_face_copy / redi_face_const (:1015-1035)
  } else if (t < a.nu + a.nv) {
    a.kv[t - a.nu] = a.ext_v ? a.ext_v[t - a.nu] : a.khtr;
  }
}

__global__ void k_face_khtr(KhtrArgs a) { face_khtr_body(a, (int)blockIdx.x); }

// ocean_redi_init fill (:493-502): Po = hEff = 0, Ko = 1
template <class F>
__global__ void k_fill_faces(size_t nrec, FaceRecF<F>* __restrict__ rec, size_t nface,
                             uint8_t* __restrict__ mono) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < nrec) rec[t] = FaceRecF<F>::put(SurfVals{0.0, 0.0, 0.0, 1, 1});
  if (t < nface) mono[t] = kMonoL | kMonoR;
}

}  // namespace redi
