// redi_kernels.hpp -- the Redi step as gfx950 kernels, in two precision forms.
//
//   k_column_state        Phase A, once per column: top-down interface
//                         pressure, PPM T/S edges and dR/dT, dR/dS (EOS).
//   k_neutral_sweep_uv    Phase A, once per wet face of both grids (a u wave
//                         and a v wave per workgroup; k_neutral_sweep for
//                         one grid): the continuous neutral surface sweep
//                         over the two adjacent columns.
//   k_tracer_recon_tiled  Phase B, per tracer group: layer averages and PPM
//                         interface values of every column.
//   k_flux (x, y)         Phase B, per tracer group and face grid: the flux of
//                         every contributing sublayer of every face, added
//                         into the bounding cells' dTr in the reference's
//                         order (W, E in x; S, N in y).
//   k_tracer_update       Phase B: hTr += dTr / areaT.
//   k_export_coeffs       face records -> the reference's Po/Ko/hEff arrays.
//
// Every kernel is a template on a precision form F:
//   Fp64  the plain form: fields, records and arithmetic in FP64, in the
//         reference's rounding order, so results match the reference bits;
//   Fp32  the FP32-store form: the fields h / hTr and the face records are
//         stored in 32-bit words (FP32 values, int16 layer indices, positions
//         as 31-bit fractions); the EOS, the PPM slopes and edges and the
//         sublayer fluxes are computed in FP32; FP64 stays in the layer
//         means and T/S interface values (the slopes and edges in FP32
//         about the means), the pressures and positions, the density and
//         guard differences, and the tendency accumulation.
// Both forms come from one source: records convert through get() / put(),
// which are identities for Fp64.
//
// Device memory layout (R = the form's storage type):
//   cell fields h, hTr          [k][j][i] (reference layout, k native bottom-up)
//   column state (K1 -> K2)     interface T, S and the next interface's
//                               pressure FP64 in both forms, interface-major
//                               within each row: ColRec of column (i, j) at
//                               interface kt (top-down, 0..nz) at
//                               (j*(nz+1) + kt)*nx + i
//   face coefficients (K2->K4)  FaceRec per face and neutral surface s,
//                               face-major: (face, s) at face*ns + s; one flag
//                               byte per face
//   tracer recon (K3 -> K4)     per column and tracer of a group, each
//                               tracer's column contiguous by default: FP64
//                               (layer mean, bottom interface value) pairs,
//                               nz+1 per column with the top interface value
//                               first; FP32-store: per layer the FP64 mean and
//                               its top and bottom values as FP32 offsets
//   dTr (K4 -> K5)              FP64, one block of nz*NT values per cell,
//                               (K, t) at (K-1)*NT + t
// The sweep reads each of its two columns once, front to back; its lanes
// are adjacent faces, which sit at similar depths, so with the column state
// interface-major per row the lanes read neighbouring records (and the
// column-state kernel, a lane per column, writes whole rows); it stages its
// face records in LDS and stores them as whole runs.  The flux kernels give
// a face one wavefront with a lane per neutral surface: a face's targets are
// non-decreasing in s, so the lanes walk down the two columns together and
// read consecutive per-column records.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "redi_column.hpp"
#include "redi_eos.hpp"

namespace redi {

// One interface k of a column (K1 output): interface T and S (FP64 in both
// forms: the sweep's density differences are differences of them), the
// pressure at the next interface down, Pn = P(min(k+1, nz)) (the surface
// pressure P(0) is 0), and dR/dT, dR/dS in the form's EOS type E: 40 bytes
// in FP64, 32 in FP32.  The sweep's window over a column holds the records
// of two interfaces and the pressures of three, the third one interface
// deeper than the records; carrying it in the record above, one load of a
// record brings both.
template <class E>
struct alignas(sizeof(E) == 8 ? 8 : 16) ColRecT {
  double T, S, Pn;
  E dT, dS;
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
template <class C>
struct LayerVals {
  C Tl, aL, aR, Tt, Tb;
};

// A tracer column (K3 output) in the plain FP64 form: per layer, its mean
// and the PPM interface value at its bottom, 16 bytes; entry 0 of a column
// holds the value at the top interface (in Tb), entry K (K = 1..nz, top-down)
// layer K.  A layer's top interface value is the bottom one of the entry
// above, since the unlimited PPM edge is one value per interface
// (redi/ocean_redi.F90:231-254), so a column stores each value once.  The
// reader gathers entries K-1 and K and runs the PPM limiter
// (redi/ocean_redi.F90:876-890) on the same three operands as the
// reference, so it obtains the same bits.
struct alignas(16) ReconPair64 {
  double Tl, Tb;
  static constexpr int kExtra = 1;  // entries per column beyond one per layer
  // layer K from entries K-1 (above) and K
  static __device__ __forceinline__ LayerVals<double> layer(const ReconPair64& above,
                                                           const ReconPair64& k) {
    double al = above.Tb, ar = k.Tb;
    (void)ppm_limit<double>(k.Tl, al, ar);
    return LayerVals<double>{k.Tl, al, ar, above.Tb, k.Tb};
  }
};

// The same layer in the FP32-store form, in 16 bytes: the FP64 layer mean
// and the interface values at its top and bottom as FP32 offsets from it.
// Differences between columns and layers are then formed as an FP64
// difference of means plus an FP32 difference of small offsets, so the
// down-gradient guard sees them to near-FP64 accuracy while the rest of
// the sublayer arithmetic runs in FP32 (the PPM limiter too, rerun on the
// offsets by the reader).
struct alignas(16) ReconRecOff {
  double Tl;
  float oT, oB;
  static constexpr int kExtra = 0;  // one entry per layer
  static __device__ __forceinline__ ReconRecOff make(double tl, double et, double eb) {
    return ReconRecOff{tl, (float)(et - tl), (float)(eb - tl)};
  }
};

// The precision forms.  R: stored fields (h, hTr); E: the EOS and its
// stored derivatives; C: the sublayer-flux arithmetic and the stored fluxes.
// In both forms the T/S interface values and the layer means are FP64 (the
// density differences that place the neutral surfaces, and the tracer
// differences the down-gradient guard tests, are differences of them), as
// are positions, pressures and the tendency accumulation.
struct Fp64 {
  using R = double;
  using E = double;
  using C = double;
  using Face = FaceRec64;
  using Col = ColRecT<double>;
  using Recon = ReconPair64;
};
struct Fp32 {
  using R = float;
  using E = float;
  using C = float;
  using Face = FaceRec32;
  using Col = ColRecT<float>;
  using Recon = ReconRecOff;
};

template <class F> using FaceRecF = typename F::Face;

// Face record (face, s) lives at face*fa + s*fb: face-major, fa = ns,
// fb = 1.  A face kernel's wave reads its face's surfaces as one contiguous
// run; the sweep, whose lanes are adjacent faces, transposes its records
// through LDS to store whole runs.
inline uint32_t face_fa(int ns) { return (uint32_t)ns; }
inline uint32_t face_fb(int) { return 1u; }

// Flag bits per face (written by the sweep): Ko of the contributing
// sublayers (hEff != 0) is non-decreasing in s on the left (bit 0) / right
// (bit 1) column; bit 2: the face has a contributing sublayer, so the flux
// kernels skip faces without one (dry faces, whose records are inert);
// bit 3: the face's records hold the inert values of a dry face (written by
// the initial fill or by the last sweep), so a sweep that finds the face dry
// again need not write them.
constexpr uint8_t kMonoL = 1, kMonoR = 2, kHasFlux = 4, kInert = 8;

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
  // (tid: the lane's index within its tile)
  __device__ __forceinline__ bool map(int bid, int nb, int& a, int& b, unsigned tid) const {
    const int tile = xcd_band(bid, nb);
    const int t0 = vertical ? tile / tiles1 : tile % tiles0;
    const int t1 = vertical ? tile % tiles1 : tile / tiles0;
    a = t0 * tx + (int)(tid % (unsigned)tx);
    b = t1 * ty + (int)(tid / (unsigned)tx);
    return a < n0 && b < n1;
  }
};

// ---------------------------------------------------------------------------
// K1: column interface state (redi_build_column, redi/ocean_redi.F90:609-654)
// ---------------------------------------------------------------------------
// The layer values and PPM edges in FP64 and the EOS derivatives in the
// form's EOS type; the interface pressure is accumulated in FP64 in both
// forms (the sweep's positions and thicknesses are differences of it).
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
  typename F::Col* rec;
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
  using C = double;
  using E = typename F::E;
  const double g_rho0 = kGravity * eos.rho0;
  const size_t cb = colstate_base((size_t)c, a.nx, nz), cs = (size_t)a.nx;
  typename F::Col* r = a.rec + cb;
  double P = 0.0;  // Pint(1) = 0
  // interface k's record is written once P(k+1) is known, at interface k+1
  C Tp = 0.0, Sp = 0.0;
  E dTp = E(0), dSp = E(0);
  constexpr bool kFp32 = std::is_same<F, Fp32>::value;
  auto load = [&](int kt, C& hh, C* s) {
    const size_t o = (size_t)(nz - 1 - kt) * n2 + c;  // top-down -> native
    hh = (C)h[o];
    const C he = hh > C(kHDivEps) ? hh : C(kHDivEps);
    if constexpr (kFp32) {  // one FP64 reciprocal per layer
      const C ih = C(1) / he;
      s[0] = (C)thtr[o] * ih;
      s[1] = (C)shtr[o] * ih;
    } else {
      s[0] = (C)thtr[o] / he;
      s[1] = (C)shtr[o] / he;
    }
  };
  auto emit = [&](int k, const C* e, const C*) {
    if (k > 0) P = P + g_rho0 * (double)h[(size_t)(nz - k) * n2 + c];  // top-down layer k-1
    E dRdT, dRdS;
    interface_drho<E>(eos, (E)e[0], (E)e[1], (E)P, dRdT, dRdS);
    if (k > 0) r[(size_t)(k - 1) * cs] = typename F::Col{Tp, Sp, P, dTp, dSp};
    Tp = e[0]; Sp = e[1]; dTp = dRdT; dSp = dRdS;
  };
  if constexpr (kFp32)  // slopes and edges in FP32 about the FP64 layer means
    stream_edges_off<2>(nz, load, emit);
  else
    stream_edges<C, 2>(nz, load, emit);
  r[(size_t)nz * cs] = typename F::Col{Tp, Sp, P, dTp, dSp};  // Pn = P(nz)
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
// (ocean_redi.F90:344-345, 367-371, 406-410), formed in FP64 from the
// stored values in both forms
template <class E>
__device__ __forceinline__ double sec_drho(const ColRecT<E>& a, const ColRecT<E>& b) {
  return 0.5 * (((double)a.dT + (double)b.dT) * (a.T - b.T) +
                ((double)a.dS + (double)b.dS) * (a.S - b.S));
}

// Sliding window over one column of the sweep.  `k` is the reference's
// kl / kr (1-based, 1..nk+1); the window holds the records at m = max(k-1,1)
// and m+1 and the interface pressures at m, m+1, min(m+2, nk+1), which is
// every index the sweep reads while the pointer sits at k.
template <class Col>
struct SweepSide {
  const Col* rec;
  size_t st;  // elements between consecutive interfaces of the column
  int k, m;
  Col w0, w1;
  double p0, p1, p2;
  __device__ __forceinline__ void init(const Col* r, size_t stride) {
    rec = r; st = stride; k = 1; m = 1;
    w0 = r[0]; w1 = r[stride];
    p0 = 0.0; p1 = w0.Pn; p2 = w1.Pn;  // P(1) = 0 at the surface, P(2), P(min(3, nk+1))
  }
  __device__ __forceinline__ Col cur() const { return k == 1 ? w0 : w1; }  // record at k
  __device__ __forceinline__ double pk() const { return k == 1 ? p0 : p1; }        // P(k)
  __device__ __forceinline__ double pk1() const { return k == 1 ? p1 : p2; }       // P(k+1)
  __device__ __forceinline__ void advance(int nk) {  // k -> k+1, k <= nk
    if (k >= 2) {
      m += 1;
      w0 = w1; w1 = rec[(size_t)m * st];
      p0 = p1; p1 = p2; p2 = w1.Pn;
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
  const typename F::Col* rec;
  FaceRecF<F>* out;        // [s][face]
  uint8_t* mono;           // [face]
  double pa_to_h;          // 1 / (g rho0): Pa -> m (ocean_redi.F90:688-691)
  const double* inv_dx;    // 1 / face spacing per face (idxCu / idyCv): the slope taper's
  double slope_max;        // opt-in slope taper: 0 = off (the reference has none)
  int keep_inert;          // 1: records of dry faces flagged kInert are left as they are
                           // (0 where the face array is not this domain's own since its fill)
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
// store fills whole 128-byte runs of records.  A stage belongs to one wave,
// and a wave's LDS instructions execute in order, so the stage needs only
// wavefront-scope fences (which order the compiler's accesses and cost no
// instruction), not workgroup barriers: the u and v waves of a workgroup
// run on without waiting for each other.
constexpr int kSweepRunWords = 8;                              // 16-byte words per face per run
// A lane's run sits kSweepLaneWords apart from the next lane's: one word of
// padding spreads the lanes' 16-byte stage writes over the LDS banks (with
// runs 128 bytes apart, every lane's write of a step lands in the same banks).
constexpr int kSweepLaneWords = kSweepRunWords + 1;
constexpr int kSweepStageWords = 64 * kSweepLaneWords;         // per wave
constexpr size_t kSweepStageBytes = kSweepStageWords * sizeof(uint4);

// `tid`: the lane's index within the tile; `stage`: the workgroup's LDS
// stage (kSweepStageWords per wave).
template <class F>
__device__ __forceinline__ void neutral_sweep_body(const SweepArgs<F>& a, const TileMap& tm,
                                                   int bid, int nb, uint4* stage, unsigned tid) {
  using Face = FaceRecF<F>;
  constexpr int kW = (int)(sizeof(Face) / sizeof(uint4));
  constexpr int kRun = kSweepRunWords / kW;  // surfaces per run
  const int lane = (int)(threadIdx.x & 63);
  uint4* stw = stage + (threadIdx.x >> 6) * kSweepStageWords;
  int fi, fj;
  const bool inr = tm.map(bid, nb, fi, fj, tid);
  const int i = a.i0 + fi, j = a.j0 + fj;
  const uint32_t fo = inr ? (uint32_t)i + (uint32_t)a.nfa * (uint32_t)j : 0u;
  const int nk = a.nz, ns = 2 * nk + 2;
  const bool wet = inr && a.wet[fo] > 0.0;  // dry face: the inert values of :722-730
  // a dry face whose records already hold them writes nothing
  const bool keep = inr && !wet && a.keep_inert && (a.mono[fo] & kInert);
  const SurfVals inert{0.0, 0.0, 0.0, 1, 1};
  auto put = [&](int s, const SurfVals& v) {
    const Face r = Face::put(v);
    uint4 w[kW];
    __builtin_memcpy(w, &r, sizeof(Face));
#pragma unroll
    for (int q = 0; q < kW; ++q) stw[lane * kSweepLaneWords + (s % kRun) * kW + q] = w[q];
  };
  // surfaces s0 .. s0+cnt-1 of every face of the wave
  auto flush = [&](int s0, int cnt) {
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    const int wd = lane % kSweepRunWords, sq = wd / kW;
#pragma unroll
    for (int it = 0; it < 64 / (64 / kSweepRunWords); ++it) {
      const int q = it * (64 / kSweepRunWords) + lane / kSweepRunWords;
      const uint32_t foq = (uint32_t)__shfl((int)fo, q, 64);
      const int okq = __shfl(inr && !keep ? 1 : 0, q, 64);
      if (okq && sq < cnt) {
        uint4* dst = reinterpret_cast<uint4*>(a.out + (size_t)foq * a.fa + (size_t)(s0 + sq) * a.fb);
        dst[wd % kW] = stw[q * kSweepLaneWords + wd];
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  };
  const size_t cl = (size_t)(i - a.di) + (size_t)a.nx * (j - a.dj);
  const size_t cr = (size_t)i + (size_t)a.nx * j;
  SweepSide<typename F::Col> L{}, Rt{};
  if (wet) {
    const size_t bl = colstate_base(cl, a.nx, nk), br = colstate_base(cr, a.nx, nk);
    L.init(a.rec + bl, (size_t)a.nx);
    Rt.init(a.rec + br, (size_t)a.nx);
  }

  int lastK_l = 1, lastK_r = 1, contK_l = 0, contK_r = 0;
  double lastP_l = 0.0, lastP_r = 0.0, lastPos_l = 0.0, lastPos_r = 0.0;
  bool reached_bottom = false, search_l = false, mono_l = true, mono_r = true, has_flux = false;
  SurfVals pend{};
  for (int ks = 1; ks <= ns; ++ks) {
   if (wet) {
    bool advL = false, advR = false;  // the side whose pointer moves on, after both branches
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
      const typename F::Col cr_ = Rt.cur();
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
        advR = true;
      } else {
        PoR = 1.0; KoR = nk;
        posR = Rt.p0 + 1.0 * (Rt.p1 - Rt.p0);
        reached_bottom = true; search_l = false;
      }
    } else {
      const typename F::Col cl_ = L.cur();
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
        advL = true;
      } else {
        PoL = 1.0; KoL = nk;
        posL = L.p0 + 1.0 * (L.p1 - L.p0);
        reached_bottom = true; search_l = true;
      }
    }
    // Moving a pointer after the two branches, not inside them, keeps its
    // record load out of the branches' register merges, which otherwise
    // wait for the load to land before the step ends.
    if (advL) L.advance(nk);
    if (advR) Rt.advance(nk);
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
        has_flux = true;
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
  if (inr)
    a.mono[fo] = wet ? (uint8_t)((mono_l ? kMonoL : 0) | (mono_r ? kMonoR : 0) | (has_flux ? kHasFlux : 0))
                     : (uint8_t)(kMonoL | kMonoR | kInert);
}

// Dynamic LDS: kSweepStageBytes per wave of the block.
template <class F>
__global__ void __launch_bounds__(256) k_neutral_sweep(SweepArgs<F> a, TileMap tm) {
  extern __shared__ uint4 sweep_stage[];
  neutral_sweep_body<F>(a, tm, (int)blockIdx.x, (int)gridDim.x, sweep_stage, threadIdx.x);
}

// Both face grids in one launch, for 64-lane tiles: wave 0 of a workgroup
// sweeps a tile of u faces, wave 1 the tile of v faces at the same tile
// position.  The u tile reads the columns of row j, the v tile those of rows
// j-1 and j, so the two waves read the columns of row j at about the same
// time, from the same caches: the column state is fetched about once for
// both grids where two launches fetch it once each.  Both maps must number
// the same tile grid (tiles0, tiles1); tiles outside a grid's faces have no
// lanes in range.  Each wave stages its records in its own part of the LDS
// and orders them with wavefront-scope fences, so neither wave waits for
// the other.
template <class F>
__global__ void __launch_bounds__(128) k_neutral_sweep_uv(SweepArgs<F> au, TileMap mu, SweepArgs<F> av,
                                                          TileMap mv) {
  extern __shared__ uint4 sweep_stage[];
  if (threadIdx.x < 64)
    neutral_sweep_body<F>(au, mu, (int)blockIdx.x, (int)gridDim.x, sweep_stage, threadIdx.x);
  else
    neutral_sweep_body<F>(av, mv, (int)blockIdx.x, (int)gridDim.x, sweep_stage, threadIdx.x - 64);
}

// ---------------------------------------------------------------------------
// K3: tracer column reconstruction (redi_tracer_column, :852-891), the NT
// tracers of a group per launch.  Entry k (0 .. ne-1, ne = nz + the form's
// kExtra) of column col and tracer t at col*ca + k*cb + t*ct (ReconLayout):
//   per column (ca = NT*ne, cb = 1, ct = ne): each tracer's column is one
//     run of entries, which a face's lanes, walking down the column, read in
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
  typename F::Recon* rec;
  ReconLayout y;
};

// Streaming form (the layer-major layout): one lane per column walks it
// top-down with a register window (stream_edges), in FP64 in both forms
// (the FP32-store form stores its edge values as FP32 offsets from the FP64
// layer means).
template <class F, int NT>
__global__ void __launch_bounds__(128) k_tracer_recon(ReconArgs<F, NT> a) {
  using C = double;
  const int N2 = a.N2, nz = a.nz;
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= N2) return;
  const size_t n2 = (size_t)N2;
  const typename F::R* __restrict__ h = a.h;
  auto load = [&](int kt, C& hh, C* s) {
    const size_t o = (size_t)(nz - 1 - kt) * n2 + c;
    hh = (C)h[o];
    const C he = hh > C(kHDivEps) ? hh : C(kHDivEps);
#pragma unroll
    for (int t = 0; t < NT; ++t) s[t] = (C)a.tr.q[t][o] / he;
  };
  C e_prev[NT];
  typename F::Recon* rc = a.rec + (size_t)c * a.y.ca;
  auto emit = [&](int k, const C* e, const C* s_above) {
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      if constexpr (F::Recon::kExtra) {  // entry k: layer k's mean (none above the top), interface k
        rc[(size_t)k * a.y.cb + (size_t)t * a.y.ct] = typename F::Recon{k > 0 ? s_above[t] : 0.0, e[t]};
      } else {
        if (k > 0)
          rc[(size_t)(k - 1) * a.y.cb + (size_t)t * a.y.ct] = F::Recon::make(s_above[t], e_prev[t], e[t]);
        e_prev[t] = e[t];
      }
    }
  };
  stream_edges<C, NT>(nz, load, emit);
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


// Row stride (doubles) of the tiled reconstruction's LDS arrays of CT
// columns.  The record stage reads them with lanes along layers, four rows
// per tracer in the FP32-store form, where a power-of-two stride puts every
// lane in the same banks: that form pads each row by one double.  The FP64
// form reads two rows per tracer and keeps the unpadded rows, whose smaller
// footprint fits one more workgroup per CU.
template <class F, int CT>
constexpr int recon_row() {
  return std::is_same<typename F::Recon, ReconRecOff>::value ? CT + 1 : CT;
}

// Element types of the tiled reconstruction's LDS stages: the layer means
// are FP64 in both forms; the thicknesses, slopes and edge offsets of the
// FP32-store form are FP32 values and are staged as such.
template <class F>
using ReconLdsSmall =
    std::conditional_t<std::is_same<typename F::Recon, ReconRecOff>::value, float, double>;

// LDS of one tiled-reconstruction workgroup, in doubles: the means, then the
// thicknesses, slopes and edges
template <class F, int NT, int KT>
constexpr int recon_lds_doubles() {
  constexpr int CTP = recon_row<F, kReconLanes / KT>();
  constexpr int small = (KT + 4) * CTP + NT * (KT + 2) * CTP + NT * (KT + 1) * CTP;
  return (KT + 4) * CTP * NT + (small * (int)sizeof(ReconLdsSmall<F>) + 7) / 8;
}
template <class F, int NT>
constexpr int recon_lds_max() {
  constexpr int a = recon_lds_doubles<F, NT, 8>(), b = recon_lds_doubles<F, NT, 16>(),
                c = recon_lds_doubles<F, NT, 32>();
  return a > b ? (a > c ? a : c) : (b > c ? b : c);
}

template <class F, int NT, int KT>
__device__ __forceinline__ void tracer_recon_tiled_body(const ReconArgs<F, NT>& a, int bid, int nb,
                                                        double* lds) {
  using R = double;
  using Sm = ReconLdsSmall<F>;  // staged thicknesses, slopes and edges
  constexpr int CT = kReconLanes / KT, NL = KT + 4, NSL = KT + 2, NE = KT + 1;
  constexpr int CTP = recon_row<F, CT>();  // row stride (elements)
  // layer values hTr / max(h, eps), layers k0-2 .. k0+KT+1 (clamped); h, same
  // layers; PLM slopes, layers k0-1 .. k0+KT; PPM edges, interfaces k0 .. k0+KT
  R(*sv)[NL][CTP] = reinterpret_cast<R(*)[NL][CTP]>(lds);
  Sm* small = reinterpret_cast<Sm*>(lds + NT * NL * CTP);
  Sm(*sh)[CTP] = reinterpret_cast<Sm(*)[CTP]>(small);
  Sm(*sd)[NSL][CTP] = reinterpret_cast<Sm(*)[NSL][CTP]>(small + NL * CTP);
  Sm(*se)[NE][CTP] = reinterpret_cast<Sm(*)[NE][CTP]>(small + NL * CTP + NT * NSL * CTP);
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
    if constexpr (std::is_same<typename F::Recon, ReconRecOff>::value) {
      // FP32-store form: FP64 means from one FP64 reciprocal per layer
      const R ih = R(1) / he;
#pragma unroll
      for (int t = 0; t < NT; ++t) sv[t][r][cc] = (R)a.tr.q[t][o] * ih;
    } else {
#pragma unroll
      for (int t = 0; t < NT; ++t) sv[t][r][cc] = (R)a.tr.q[t][o] / he;
    }
  }
  __syncthreads();
  // The FP32-store form (ReconRecOff records) runs stages 2-4 in FP32 on
  // differences of the FP64 layer means: PLM slopes and PPM edges use only
  // differences of their layer values (plus, for an edge, one of them), so
  // shifting every value by one layer's mean changes them by exactly that
  // shift; each slope is formed about its own layer's mean, each edge about
  // the mean of the layer above it, and the records keep the edges as
  // offsets from their layer's mean.
  constexpr bool kOff = std::is_same<typename F::Recon, ReconRecOff>::value;
  using W = std::conditional_t<kOff, float, R>;  // arithmetic of stages 2-3
  // 2: slopes of layers kl = k0-1+r; 0 for the first and last layer (the
  // slope stream_edges starts from and the one it skips)
  for (int q = tid; q < NSL * CT; q += kReconLanes) {
    const int r = q / CT, cc = q % CT;
    const int kl = k0 - 1 + r;  // rows r, r+1, r+2 of sh / sv hold layers kl-1, kl, kl+1
    const bool in = kl >= 1 && kl <= nz - 2;
    // thickness weights once for every tracer of the group
    const PlmW<W> w = plm_weights<W>((W)sh[r][cc], (W)sh[r + 1][cc], (W)sh[r + 2][cc]);
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      if constexpr (kOff) {
        sd[t][r][cc] = in ? (R)plm_slope_w<float>(w, (float)(sv[t][r][cc] - sv[t][r + 1][cc]), 0.f,
                                                  (float)(sv[t][r + 2][cc] - sv[t][r + 1][cc]))
                          : R(0);
      } else {
        sd[t][r][cc] = in ? plm_slope_w<R>(w, sv[t][r][cc], sv[t][r + 1][cc], sv[t][r + 2][cc]) : R(0);
      }
    }
  }
  __syncthreads();
  // 3: edges at interfaces ki = k0+r
  for (int q = tid; q < NE * CT; q += kReconLanes) {
    const int r = q / CT, cc = q % CT;
    const int ki = k0 + r;  // sh / sv rows r..r+3: layers ki-2 .. ki+1 (clamped);
                            // sd rows r, r+1: slopes of layers ki-1, ki
    const PpmW<W> w = ppm_weights<W>((W)sh[r][cc], (W)sh[r + 1][cc], (W)sh[r + 2][cc], (W)sh[r + 3][cc]);
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      R e = R(0);
      if constexpr (kOff) {
        // offset from the mean of layer ki-1 (ki = 0 and ki = nz: 0, the edge
        // being the adjacent layer's mean there)
        if (ki > 0 && ki < nz)
          e = (R)ppm_edge_w<float>(w, 0.f, (float)(sv[t][r + 2][cc] - sv[t][r + 1][cc]),
                                   (float)sd[t][r][cc], (float)sd[t][r + 1][cc]);
      } else if (ki == 0) {
        e = sv[t][r + 2][cc] - R(0.5) * R(0);  // edge(1) = tr(1) - 0.5*diff(1), diff(1) = 0
      } else if (ki == nz) {
        e = sv[t][r + 1][cc] + R(0.5) * R(0);  // diff(nk) = 0
      } else if (ki < nz) {
        e = ppm_edge_w<R>(w, sv[t][r + 1][cc], sv[t][r + 2][cc], sd[t][r][cc], sd[t][r + 1][cc]);
      }
      se[t][r][cc] = e;
    }
  }
  __syncthreads();
  // 4: records of layers kl = k0+kk, lanes along layers
  const int kk = tid % KT, cc = tid / KT;
  const int kl = k0 + kk, c = c0 + cc;
  if (kl < nz && c < N2) {
    typename F::Recon* rc = a.rec + (size_t)c * a.y.ca + (size_t)kl * a.y.cb;
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      if constexpr (kOff) {
        // top edge: offset from layer kl-1's mean, moved to layer kl's; bottom
        // edge: already about layer kl's mean
        const float oT = kl == 0 ? 0.f : (float)((sv[t][kk + 1][cc] - sv[t][kk + 2][cc]) + se[t][kk][cc]);
        rc[(size_t)t * a.y.ct] = ReconRecOff{sv[t][kk + 2][cc], oT, (float)se[t][kk + 1][cc]};
      } else {
        // entry kl + 1: layer kl's mean and bottom value; entry 0: the top value
        rc[(size_t)a.y.cb + (size_t)t * a.y.ct] = typename F::Recon{sv[t][kk + 2][cc], se[t][kk + 1][cc]};
        if (kl == 0) rc[(size_t)t * a.y.ct] = typename F::Recon{0.0, se[t][kk][cc]};
      }
    }
  }
}

template <class F, int NT, int KT>
__global__ void __launch_bounds__(kReconLanes) k_tracer_recon_tiled(ReconArgs<F, NT> a) {
  __shared__ double lds[recon_lds_doubles<F, NT, KT>()];
  tracer_recon_tiled_body<F, NT, KT>(a, (int)blockIdx.x, (int)gridDim.x, lds);
}

// Workgroups of the tiled reconstruction of one domain.
inline unsigned recon_tiled_blocks(int N2, int nz, int kt) {
  const int ct = kReconLanes / kt;
  return (unsigned)((N2 + ct - 1) / ct) * (unsigned)((nz + kt - 1) / kt);
}

// ---------------------------------------------------------------------------
// K4: sublayer fluxes of one face grid and their divergence into the cells
// (redi_face_flux + redi_sublayer_dT, :893-925, :1037-1092, accumulated as
// redi_apply_flux_impl does, :1094-1181)
// ---------------------------------------------------------------------------

// Pieces of redi_sublayer_dT (:893-925).  The interface value at position
// Po in layer K of a column, (1 - Po) * T(K-1) + Po * T(K), from the values
// at the layer's top (tt) and bottom (tb) interfaces (the tlt / tlb / trt /
// trb expressions, :905-912):
template <class C>
__device__ __forceinline__ C edge_value(C tt, C tb, C Po) {
  return (C(1) - Po) * tt + Po * tb;
}

// The down-gradient guard (:913-924): the flux is zeroed unless the top
// and bottom differences agree in sign and the mean and layer differences
// agree in sign.
template <class R>
__device__ __forceinline__ bool guard_zero(R dT_top, R dT_bot, R dT_ave, R dT_layer) {
  return sgn_prod_le0<R>(dT_top, dT_bot) || sgn_prod_le0<R>(dT_ave, dT_layer);
}

// redi_sublayer_dT's result from the interface values at the sublayer's
// top (tlt, trt) and bottom (tlb, trb) and the layer means: the layer
// difference, or 0 where the guard zeroes it (`zero`).
template <class R>
__device__ __forceinline__ R guarded_dT(R tlt, R tlb, R trt, R trb, R tlay, R trlay, bool& zero) {
  const R dT_top = trt - tlt;
  const R dT_bot = trb - tlb;
  const R dT_ave = R(0.5) * (dT_top + dT_bot);
  const R dT_layer = trlay - tlay;
  zero = guard_zero<R>(dT_top, dT_bot, dT_ave, dT_layer);
  return zero ? R(0) : dT_layer;
}

constexpr int kMaxNT = 4;

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

// The reference adds, for every cell, the fluxes of its four faces into
// dTr(1:nz) in a fixed order -- W, E, S, N, each face's sublayers
// ascending (:1136-1175) -- and FP64 addition is not associative, so that
// order is part of the result.  Here the order is kept by walking the
// faces in it:
//
//   k_flux_x  one wavefront per run of `seg` cells along a row.  It takes
//             the row's u faces left to right; face i's right-side fluxes
//             are cell i's W contributions (its first, added to 0) and its
//             left-side fluxes are cell i-1's E contributions (its second),
//             after which cell i-1's partial dTr is stored.
//   k_flux_y  one wavefront per run of cells up a column, the column's v
//             faces south to north: face j's right-side fluxes are cell
//             j's S contributions (added to its stored W+E partial), its
//             left-side fluxes cell j-1's N contributions (the last), after
//             which cell j-1's dTr is complete and stored.
//   k_tracer_update  hTr += dTr * (1/areaT), per cell and layer.
//
// A run's first face (whose left cell belongs to the previous run) and last
// face (whose right cell belongs to the next) are computed by both runs, the
// same expression on the same operands, so both get the same bits.
//
// Within a face, one lane per neutral surface computes the sublayer fluxes
// (every sublayer's flux once, where the reference evaluates it from each
// bounding cell: the same expression on the same operands); the
// contributing ones (hEff != 0) are compacted into LDS in s order with
// their two target layers.  Then one lane per (layer K, tracer) of the
// cell adds its layer's contributions in s order.  Along a face the
// targets of each side normally never decrease, so layer K's contributions
// are one run of the compacted list, found from a per-layer start table;
// a side whose targets decrease somewhere is scanned whole by every lane
// instead, which is the same additions in the same order.
//
// dTr per cell lives in HBM between the two kernels (FP64, `acc`, the
// (cell, K, t) entry at cell*nz*NT + (K-1)*NT + t), one contiguous block per
// cell, so each lane's loads and stores are coalesced.
constexpr int kChunk = 63;      // sublayers per pass of a wavefront (64 surfaces)
constexpr int kFluxSegDefault = 16;  // cells per run

// Layer-capacity classes: a domain of nz layers runs the kernels of the
// smallest class with nzc >= nz (the LDS lists and the register
// accumulators are sized by the class).
constexpr int kFluxNzc[4] = {32, 76, 128, 256};

template <class F>
struct FluxArgs {
  int nx, ny, nz;         // cells per row, rows, layers
  int wall0, wall1;       // 0-based face coordinate across the faces (i for u,
                          // j for v) of the two wall faces, -1 where open
  int seg;                // cells per run
  int nseg;               // runs per line of cells (row for x, column for y)
  uint32_t fa, fb;        // face record (face, s) at face*fa + s*fb
  double dt;
  const double *khtr, *len, *inv;  // face diffusivity, face length, 1 / spacing
  const FaceRecF<F>* rec;
  const uint8_t* flags;   // face flags (kHasFlux)
  ReconLayout ry;
  const typename F::Recon* recon;  // tracer t's records from recon + t*ry.ct
  double* acc;            // dTr, (cell, K, t) at cell*nz*NT + (K-1)*NT + t
  // optional diagnostic: for tracer t of the group, face fo and chunk c of
  // 63 sublayers, the bits of the contributing sublayers the guard zeroed,
  // at guard[t*guard_ts + fo*guard_fs + c] (null: not written)
  uint64_t* guard;
  size_t guard_ts, guard_fs;
  int guard_nt;           // tracers of the group that have a block (a lone tracer's copy has none)
};

// LDS of one flux wavefront: the compacted contributions of one face
// (fluxes in the form's compute type).
template <class X, int NT, int NZC>
struct FluxLds {
  X x[(2 * NZC + 1) * NT];           // fluxes, entry e's tracer t at e*NT + t
  uint16_t tgt[2][2 * NZC + 1];      // entry e's target layer, left / right column
  uint16_t start[2][NZC + 2];        // first entry whose target >= K (K = 1..nz+1)
};

// The contributing sublayers of face `fo` between columns cl (left) and cr
// (right), compacted into L in s order; returns their count (wave-uniform).
// Two ways of keeping loads in flight, each where the registers allow four
// waves per SIMD (kY: the v-face kernel, which also holds the cell's W + E
// partial):
//   kAhead (FP64 form, u faces): each tracer's reconstruction entries are
//     gathered while the previous tracer is computed (16 VGPRs);
//   kPre (FP64 form; FP32-store form, v faces): the next chunk's surface
//     records are loaded while this chunk is computed (8 VGPRs; the
//     FP32-store u-face kernel would drop from five waves per SIMD to four,
//     which costs more than it hides).
template <class F, int NT, int NZC, bool kY>
__device__ __forceinline__ int face_contributions(const FluxArgs<F>& a, size_t fo, uint32_t cl,
                                                  uint32_t cr, FluxLds<typename F::C, NT, NZC>& L) {
  using C = typename F::C;
  using Face = FaceRecF<F>;
  constexpr bool kOff = std::is_same<typename F::Recon, ReconRecOff>::value;
  constexpr bool kAhead = !kOff && !kY;
  constexpr bool kPre = !kOff || kY;
  const int lane = (int)(threadIdx.x & 63);
  const C coef = (C)(((a.dt * a.khtr[fo]) * a.len[fo]) * a.inv[fo]);
  const typename F::Recon* recL = a.recon + cl * a.ry.ca;
  const typename F::Recon* recR = a.recon + cr * a.ry.ca;
  const uint32_t cb = a.ry.cb, ct = a.ry.ct;
  const size_t fb = a.fb;
  const int ns = 2 * a.nz + 2, nsub = ns - 1;
  const Face* r = a.rec + fo * a.fa;
  int nc = 0;
  Face rn{};  // kPre: the surface records of the chunk at hand
  if (kPre && lane < ns) rn = r[(size_t)lane * fb];
  for (int c0 = 0; c0 < nsub; c0 += kChunk) {
    const int s = c0 + lane;
    const bool surf = s < ns;
    SurfVals Fs{0.0, 0.0, 0.0, 1, 1};
    if constexpr (kPre) {
      const Face rc = rn;
      if (s + kChunk < ns) rn = r[(size_t)(s + kChunk) * fb];  // the next chunk's
      if (surf) Fs = rc.get();
    } else {
      if (surf) Fs = r[(size_t)s * fb].get();
    }
    const int KoL = Fs.KoL, KoR = Fs.KoR;
    const C PoL = (C)Fs.PoL, PoR = (C)Fs.PoR;
    // surface s+1, the bottom of sublayer s (lane 63's value is never used)
    const C PoLb = __shfl_down(PoL, 1, 64), PoRb = __shfl_down(PoR, 1, 64);
    const int KoLb = __shfl_down(KoL, 1, 64), KoRb = __shfl_down(KoR, 1, 64);
    const bool contrib = surf && lane < kChunk && s < nsub && Fs.hEff != 0.0;
    const C hEff = (C)Fs.hEff;
    C x[NT];
    // The FP32-store form gathers every tracer's two 16-byte records before
    // computing (one round trip per chunk, eight records in flight); the FP64
    // form gathers one tracer's four 16-byte entries at a time (kAhead: the
    // next tracer's while this one is computed), which keeps its registers,
    // and so its waves per SIMD, at the faster setting.
    ReconRecOff gl[kOff ? NT : 1], gr[kOff ? NT : 1];
    // FP64: entries K-1 and K of both columns (with kAhead, one tracer ahead)
    ReconPair64 pla[2], plk[2], pra[2], prk[2];
    auto gather64 = [&](int t, int b) {
      if constexpr (!kOff) {
        pla[b] = plk[b] = pra[b] = prk[b] = ReconPair64{0.0, 0.0};
        if (surf) {
          pla[b] = recL[(uint32_t)(KoL - 1) * cb + (uint32_t)t * ct];
          plk[b] = recL[(uint32_t)KoL * cb + (uint32_t)t * ct];
          pra[b] = recR[(uint32_t)(KoR - 1) * cb + (uint32_t)t * ct];
          prk[b] = recR[(uint32_t)KoR * cb + (uint32_t)t * ct];
        }
      }
    };
    if constexpr (kAhead) gather64(0, 0);
    if constexpr (kOff) {
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        gl[t] = gr[t] = ReconRecOff{0.0, 0.f, 0.f};
        if (surf) {
          gl[t] = recL[(uint32_t)(KoL - 1) * cb + (uint32_t)t * ct];
          gr[t] = recR[(uint32_t)(KoR - 1) * cb + (uint32_t)t * ct];
        }
      }
    }
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      x[t] = C(0);
      bool zero = false;
      if constexpr (kOff) {
        // FP32-store form: FP64 means, FP32 offsets.  The sublayer's top
        // difference is D + (eR - eL) with D the FP64 difference of the two
        // layer means and eL, eR FP32 edge offsets; its bottom difference is
        // the next surface's top one; its layer difference is D plus the
        // difference of the FP32 parabola means about the layer means.
        const ReconRecOff ql = gl[t], qr = gr[t];
        const double D = qr.Tl - ql.Tl;
        const float eL = (1.f - PoL) * ql.oT + PoL * ql.oB;
        const float eR = (1.f - PoR) * qr.oT + PoR * qr.oB;
        const double dtop = D + (double)(eR - eL);
        const double dbot = __shfl_down(dtop, 1, 64);
        if (contrib) {
          float aLl = ql.oT, aRl = ql.oB, aLr = qr.oT, aRr = qr.oB;
          (void)ppm_limit<float>(0.f, aLl, aRl);
          (void)ppm_limit<float>(0.f, aLr, aRr);
          const float pl = ppm_ave<float>(PoL, PoLb + (float)(KoLb - KoL), aLl, aRl, 0.f);
          const float pr = ppm_ave<float>(PoR, PoRb + (float)(KoRb - KoR), aLr, aRr, 0.f);
          const double dlay = D + (double)(pr - pl);
          const double dave = 0.5 * (dtop + dbot);
          zero = guard_zero<double>(dtop, dbot, dave, dlay);
          x[t] = zero ? C(0) : ((C)dlay * hEff) * coef;
        }
      } else {
        // entries K-1 and K of each column (layer K's top value, mean and
        // bottom value); the limited edges only where a flux is formed
        if constexpr (kAhead) {
          if (t + 1 < NT) gather64(t + 1, (t + 1) & 1);
        } else {
          gather64(t, t & 1);
        }
        const ReconPair64 la = pla[t & 1], lk = plk[t & 1], ra = pra[t & 1], rk = prk[t & 1];
        const C eL = edge_value<C>(la.Tb, lk.Tb, PoL), eR = edge_value<C>(ra.Tb, rk.Tb, PoR);
        const C eLb = __shfl_down(eL, 1, 64), eRb = __shfl_down(eR, 1, 64);
        if (contrib) {
          const LayerVals<C> qL = ReconPair64::layer(la, lk), qR = ReconPair64::layer(ra, rk);
          const C tlay = ppm_ave<C>(PoL, PoLb + (C)(KoLb - KoL), qL.aL, qL.aR, qL.Tl);
          const C trlay = ppm_ave<C>(PoR, PoRb + (C)(KoRb - KoR), qR.aL, qR.aR, qR.Tl);
          const C d = guarded_dT<C>(eL, eLb, eR, eRb, tlay, trlay, zero);
          x[t] = (d * hEff) * coef;
        }
      }
      if (a.guard && t < a.guard_nt) {  // wave-uniform
        const uint64_t gm = __ballot(contrib && zero);
        if (lane == 0) a.guard[(size_t)t * a.guard_ts + fo * a.guard_fs + (size_t)(c0 / kChunk)] = gm;
      }
    }
    const uint64_t m = __ballot(contrib);
    const int pos = nc + (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                                         __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
    if (contrib) {
#pragma unroll
      for (int t = 0; t < NT; ++t) L.x[pos * NT + t] = x[t];
      L.tgt[0][pos] = (uint16_t)KoL;
      L.tgt[1][pos] = (uint16_t)KoR;
    }
    nc += __popcll(m);
  }
  return nc;
}

// Lane `lane` of pass p takes entry idx = p*64 + lane of a cell's dTr block,
// layer K = idx / NT + 1 and tracer t = idx % NT.
//
// v + (sign) the contributions of one side (0 left, 1 right) of the face in
// L to layer K, tracer t, in s order.
template <class X, int NT, int NZC>
__device__ __forceinline__ double fold_side(double v, const FluxLds<X, NT, NZC>& L, int side, bool mono,
                                            int nc, int K, int t, bool plus) {
  if (mono) {
    const int lo = L.start[side][K], hi = L.start[side][K + 1];
    for (int e = lo; e < hi; ++e) {
      const double x = (double)L.x[e * NT + t];
      v = plus ? v + x : v - x;
    }
  } else {
    for (int e = 0; e < nc; ++e)
      if ((int)L.tgt[side][e] == K) {
        const double x = (double)L.x[e * NT + t];
        v = plus ? v + x : v - x;
      }
  }
  return v;
}

// Run `w` of the x (kY = false) or y (kY = true) face grid.
template <class F, int NT, int NZC, bool kY>
__device__ __forceinline__ void flux_body(const FluxArgs<F>& a, int bid, int nb,
                                          FluxLds<typename F::C, NT, NZC>& L) {
  constexpr int P = (NZC * NT + 63) / 64;  // passes over a cell's dTr block
  const int lane = (int)(threadIdx.x & 63);
  const int w = xcd_band(bid, nb);
  const int nx = a.nx, ny = a.ny, nz = a.nz, nk = nz * NT;
  // x: runs along row `line`; y: runs up column `line`, the runs of one
  // height side by side (adjacent columns share their records' cache lines)
  const int line = kY ? w % nx : w / a.nseg;
  const int sg = kY ? w / nx : w % a.nseg;
  const int n_along = kY ? ny : nx;
  const int c_lo = sg * a.seg;
  if (line >= (kY ? nx : ny) || c_lo >= n_along) return;
  const int c_hi = c_lo + a.seg < n_along ? c_lo + a.seg : n_along;
  Global<double>* acc = global_ptr(a.acc);
  // dTr partials of the run's current cells: prev = the left cell of the
  // face at hand, cur = its right cell
  double prev[P], cur[P];
#pragma unroll
  for (int p = 0; p < P; ++p) prev[p] = cur[p] = 0.0;
  for (int f = c_lo; f <= c_hi; ++f) {
    const bool doL = f > c_lo;  // the face's left cell is this run's (E / N)
    const bool doR = f < c_hi;  // its right cell is this run's (W / S)
    // right column (i, j) of the face, its left column (i - 1, j) or (i, j - 1)
    const int i = kY ? line : f, j = kY ? f : line;
    const uint32_t cr = (uint32_t)i + (uint32_t)nx * (uint32_t)j;
    const uint32_t cl = kY ? cr - (uint32_t)nx : cr - 1u;
    if (kY && doR) {  // the cell's W + E partial, loaded early
      const Global<double>* src = acc + (size_t)cr * nk;
#pragma unroll
      for (int p = 0; p < P; ++p) {
        const int idx = p * 64 + lane;
        cur[p] = idx < nk ? src[idx] : 0.0;
      }
    }
    // faces computed: i = 2..nx (u), j = 2..ny (v), 1-based, and not a wall
    const int across = kY ? j : i;
    const size_t fo = kY ? (size_t)i + (size_t)nx * j : (size_t)i + (size_t)(nx + 1) * j;
    const bool live = across >= 1 && across <= n_along - 1 && across != a.wall0 &&
                      across != a.wall1 && (a.flags[fo] & kHasFlux);
    int nc = 0;
    bool monoL = true, monoR = true;
    if (live) {
      nc = face_contributions<F, NT, NZC, kY>(a, fo, cl, cr, L);
      __syncthreads();
      // a side is monotone when no entry's target is below the previous one's
      bool dL = false, dR = false;
      for (int e = lane + 1; e < nc; e += 64) {
        dL = dL || L.tgt[0][e] < L.tgt[0][e - 1];
        dR = dR || L.tgt[1][e] < L.tgt[1][e - 1];
      }
      monoL = __ballot(dL) == 0;
      monoR = __ballot(dR) == 0;
      for (int k = lane; k < nz + 2; k += 64) {
        L.start[0][k] = (uint16_t)nc;
        L.start[1][k] = (uint16_t)nc;
      }
      __syncthreads();
      // start[K] = first entry with target >= K: entry e starts layers
      // target(e-1)+1 .. target(e)
      for (int e = lane; e < nc; e += 64) {
        if (monoL) {
          const int hi = L.tgt[0][e], lo = e ? L.tgt[0][e - 1] + 1 : 1;
          for (int k = lo; k <= hi; ++k) L.start[0][k] = (uint16_t)e;
        }
        if (monoR) {
          const int hi = L.tgt[1][e], lo = e ? L.tgt[1][e - 1] + 1 : 1;
          for (int k = lo; k <= hi; ++k) L.start[1][k] = (uint16_t)e;
        }
      }
      __syncthreads();
    }
#pragma unroll
    for (int p = 0; p < P; ++p) {
      const int idx = p * 64 + lane;
      const int K = idx / NT + 1, t = idx % NT;
      const bool ok = idx < nk;
      if (doL) {  // E (x) / N (y): the left cell's second / last face
        if (ok && nc) prev[p] = fold_side(prev[p], L, 0, monoL, nc, K, t, true);
        if (ok) acc[(size_t)cl * nk + idx] = prev[p];
      }
      if (doR) {  // W (x, added to 0) / S (y, added to the W + E partial)
        double v = kY ? cur[p] : 0.0;
        if (ok && nc) v = fold_side(v, L, 1, monoR, nc, K, t, false);
        prev[p] = v;
      }
    }
    if (live) __syncthreads();  // the next face overwrites L
  }
}

template <class F, int NT, int NZC, bool kY>
__global__ void __launch_bounds__(64) k_flux(FluxArgs<F> a) {
  __shared__ FluxLds<typename F::C, NT, NZC> lds;
  flux_body<F, NT, NZC, kY>(a, (int)blockIdx.x, (int)gridDim.x, lds);
}

// Runs (workgroups) of one face grid.
inline unsigned flux_blocks(int nx, int ny, int seg, bool y, int* nseg) {
  const int n = y ? ny : nx;
  *nseg = (n + seg - 1) / seg;
  return (unsigned)*nseg * (unsigned)(y ? nx : ny);
}

// ---------------------------------------------------------------------------
// K5: hTr += dTr * (1/areaT) (redi_apply_flux_impl, :1177-1179)
// ---------------------------------------------------------------------------
template <class F>
struct UpdateArgs {
  int N2, nz, nt;                   // columns, layers, tracers of the group
  const double* areaT;
  const double* acc;                // dTr, the flux kernels' layout
  typename F::R* htr[kMaxNT];
  double* tend[kMaxNT];             // optional: the FP64 tendency dTr / areaT per cell
};

// A workgroup takes kUpdCells consecutive columns x kUpdK layers: it reads
// their dTr blocks along each column's run of (K, t) entries, transposes
// them through LDS, and updates each tracer layer as one contiguous run of
// columns.
constexpr int kUpdCells = 64, kUpdK = 8, kUpdLanes = 256;

template <class F>
__device__ __forceinline__ void tracer_update_body(const UpdateArgs<F>& a, int bid, int nb,
                                                   double* tile, double* inv) {
  using R = typename F::R;
  const int nz = a.nz, NT = a.nt, N2 = a.N2, nk = nz * NT;
  const int nkb = (nz + kUpdK - 1) / kUpdK;
  const int b = xcd_band(bid, nb);
  const int c0 = (b / nkb) * kUpdCells, k0 = (b % nkb) * kUpdK;  // k0: 0-based top-down layer
  const int rows = (nz - k0 < kUpdK ? nz - k0 : kUpdK) * NT;      // (K, t) rows of the tile
  const int tid = (int)threadIdx.x;
  constexpr int kRowMax = kUpdK * kMaxNT;
  const Global<const double>* acc = global_ptr(a.acc);
  for (int q = tid; q < kUpdCells * kRowMax; q += kUpdLanes) {
    const int cc = q / kRowMax, r = q % kRowMax, c = c0 + cc;
    if (c < N2 && r < rows) tile[r * (kUpdCells + 1) + cc] = acc[(size_t)c * nk + (size_t)k0 * NT + r];
  }
  if (tid < kUpdCells && c0 + tid < N2) inv[tid] = 1.0 / global_ptr(a.areaT)[c0 + tid];
  __syncthreads();
  for (int q = tid; q < kRowMax * kUpdCells; q += kUpdLanes) {
    const int r = q / kUpdCells, cc = q % kUpdCells, c = c0 + cc;
    if (c >= N2 || r >= rows) continue;
    const int k = k0 + r / NT, t = r % NT;
    R* hq = a.htr[0];
    double* tq = a.tend[0];
#pragma unroll
    for (int u = 1; u < kMaxNT; ++u)
      if (t == u) { hq = a.htr[u]; tq = a.tend[u]; }
    const size_t o = (size_t)(nz - 1 - k) * (size_t)N2 + c;  // top-down layer k+1 -> native
    const double d = tile[r * (kUpdCells + 1) + cc] * inv[cc];
    Global<R>* hp = global_ptr(hq);
    hp[o] = (R)((double)hp[o] + d);
    if (tq) global_ptr(tq)[o] = d;
  }
}

constexpr int kUpdTileDoubles = kUpdK * kMaxNT * (kUpdCells + 1);

template <class F>
__global__ void __launch_bounds__(kUpdLanes) k_tracer_update(UpdateArgs<F> a) {
  __shared__ double tile[kUpdTileDoubles];
  __shared__ double inv[kUpdCells];
  tracer_update_body<F>(a, (int)blockIdx.x, (int)gridDim.x, tile, inv);
}

inline unsigned update_blocks(int N2, int nz) {
  return (unsigned)((N2 + kUpdCells - 1) / kUpdCells) * (unsigned)((nz + kUpdK - 1) / kUpdK);
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
  if (t < nface) mono[t] = kMonoL | kMonoR | kInert;
}

}  // namespace redi
