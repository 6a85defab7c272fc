// redi_lib.hip -- the C ABI of include/redi_hip.h over the gfx950 kernels.
//
// Host-side restatement of the reference's outer shims: ocean_redi_init /
// destroy (redi/ocean_redi.F90:466-529), redi_calc_coeffs (:572-607) and
// redi_apply_flux (:927-1002) with the same silent no-op guards, plus the
// flat launchers with the argument list of redi/opt_kernel.cu:534-574.
#include <hip/hip_runtime.h>
#include <rocprofiler-sdk-roctx/roctx.h>
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/redi_hip.h"
#include "redi_kernels.hpp"

using namespace redi;

namespace {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

// A named host-side range around a public call, from its entry to its
// return (rocprofv3 --marker-trace shows it beside the kernels it
// enqueued; without a profiler attached the calls do nothing).
struct Range {
  explicit Range(const char* name) { roctxRangePushA(name); }
  ~Range() { roctxRangePop(); }
  Range(const Range&) = delete;
  Range& operator=(const Range&) = delete;
};

#define HIP_TRY(expr)                                                              \
  do {                                                                             \
    hipError_t e_ = (expr);                                                        \
    if (e_ != hipSuccess)                                                          \
      return fail(REDI_ERR_HIP, "%s failed: %s", #expr, hipGetErrorString(e_));    \
  } while (0)

inline unsigned blocks_for(size_t n, unsigned bs) { return (unsigned)((n + bs - 1) / bs); }

// HIP-event brackets around kernel launches (see redi_timing_enable).
struct Timer {
  bool on = false;
  struct Rec { int kind; hipEvent_t a, b; };
  std::vector<Rec> live;
  std::vector<hipEvent_t> pool;
  hipEvent_t get() {
    if (!pool.empty()) { hipEvent_t e = pool.back(); pool.pop_back(); return e; }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
  }
  int begin(int kind, hipStream_t st) {
    if (!on) return -1;
    Rec r{kind, get(), get()};
    if (!r.a || !r.b || hipEventRecord(r.a, st) != hipSuccess) return -1;
    live.push_back(r);
    return (int)live.size() - 1;
  }
  void end(int slot, hipStream_t st) {
    if (slot >= 0) (void)hipEventRecord(live[slot].b, st);
  }
  ~Timer() {
    for (auto& r : live) { (void)hipEventDestroy(r.a); (void)hipEventDestroy(r.b); }
    for (auto e : pool) (void)hipEventDestroy(e);
  }
};
Timer g_no_timer;

constexpr int kTracerGroup = kMaxNT;  // tracers per flux launch

// Tracer slots of the Phase-B scratch for an nt-tracer state (a lone
// tracer runs through the two-tracer kernels, so at least two).
int recon_slots(int nt) { return std::min(std::max(nt, 2), kTracerGroup); }

// Faces of a face-record array: the face count rounded up to a multiple of
// 4, so every face grid's records start 256-byte aligned.
constexpr size_t kFacePad = 4;
size_t padded_faces(size_t nf) { return (nf + kFacePad - 1) / kFacePad * kFacePad; }

// Bytes of the per-form records and values (the column state is FP64 in
// both forms).
struct FormSizes {
  size_t face, recon, val, col;  // FaceRec, reconstruction entry, stored value, ColRec
  int recon_extra;               // reconstruction entries per column beyond one per layer
};
template <class F>
constexpr FormSizes form_sizes() {
  return FormSizes{sizeof(FaceRecF<F>), sizeof(typename F::Recon), sizeof(typename F::R),
                   sizeof(typename F::Col), F::Recon::kExtra};
}
FormSizes sizes_of(int form) {
  return form == REDI_FORM_FP32_STORE ? form_sizes<Fp32>() : form_sizes<Fp64>();
}

// Device scratch of one step; Phase A and Phase B are live at different
// times and share one buffer.
//   Phase A: a ColRec per column interface.
//   Phase B, per tracer slot: the reconstruction entries of every column
//   (one per layer, plus recon_extra) and an FP64 dTr value per layer;
//   once: a discard layer field that takes the second output of a
//   two-tracer launch run for a lone tracer.
// Every array starts on a 256-byte boundary.
struct PhaseBScratch {
  void* recon;       // N2 * (nz + recon_extra) * slots entries (ReconLayout)
  void* discard;     // N2 * nz values
  double* acc;       // N2 * nz * slots dTr values
  size_t bytes;
};

PhaseBScratch carve_phase_b(void* base, int nx, int ny, int nz, int slots, const FormSizes& z) {
  const size_t N2 = (size_t)nx * ny;
  char* b = (char*)base;
  size_t off = 0;
  auto take = [&](size_t bytes) {
    char* p = b ? b + off : nullptr;
    off += (bytes + 255) & ~(size_t)255;
    return p;
  };
  PhaseBScratch s;
  s.recon = take(z.recon * N2 * (size_t)(nz + z.recon_extra) * slots);
  s.discard = take(z.val * N2 * nz);
  s.acc = (double*)take(sizeof(double) * N2 * nz * slots);
  s.bytes = off;
  return s;
}

size_t workspace_bytes(int nx, int ny, int nz, int nt, int form) {
  const FormSizes z = sizes_of(form);
  const size_t phase_a = z.col * (nz + 1) * (size_t)nx * ny;
  return std::max(phase_a, carve_phase_b(nullptr, nx, ny, nz, recon_slots(nt), z).bytes);
}

// The per-face Phase-A state: FaceRec (face, s) at face*ns + s
// (face-major, the face count padded) for u and v faces, plus the face
// flags and the resolved face diffusivities.
struct FaceState {
  void* u = nullptr;
  void* v = nullptr;
  uint8_t* mono_u = nullptr;
  uint8_t* mono_v = nullptr;
  double* khtr_u = nullptr;
  double* khtr_v = nullptr;
  size_t nu = 0, nv = 0;    // faces
  size_t nup = 0, nvp = 0;  // faces per surface row
  bool filled = false;      // filled for this domain (fill_face_state), so kInert flags hold
};

size_t face_state_bytes(int nx, int ny, int nz, int form) {
  const size_t ns = 2 * (size_t)nz + 2, nu = (size_t)(nx + 1) * ny, nv = (size_t)nx * (ny + 1);
  return (padded_faces(nu) + padded_faces(nv)) * ns * sizes_of(form).face +
         (nu + nv) * (1 + sizeof(double));
}

int alloc_face_state(FaceState* fs, int nx, int ny, int nz, bool own_khtr, int form) {
  const size_t ns = 2 * (size_t)nz + 2, rb = sizes_of(form).face;
  fs->nu = (size_t)(nx + 1) * ny;
  fs->nv = (size_t)nx * (ny + 1);
  fs->nup = padded_faces(fs->nu);
  fs->nvp = padded_faces(fs->nv);
  HIP_TRY(hipMalloc(&fs->u, fs->nup * ns * rb));
  HIP_TRY(hipMalloc(&fs->v, fs->nvp * ns * rb));
  HIP_TRY(hipMalloc(&fs->mono_u, fs->nu));
  HIP_TRY(hipMalloc(&fs->mono_v, fs->nv));
  if (own_khtr) {
    HIP_TRY(hipMalloc(&fs->khtr_u, fs->nu * sizeof(double)));
    HIP_TRY(hipMalloc(&fs->khtr_v, fs->nv * sizeof(double)));
  }
  return REDI_OK;
}

void free_face_state(FaceState* fs, bool own_khtr) {
  (void)hipFree(fs->u); (void)hipFree(fs->v);
  (void)hipFree(fs->mono_u); (void)hipFree(fs->mono_v);
  if (own_khtr) { (void)hipFree(fs->khtr_u); (void)hipFree(fs->khtr_v); }
  *fs = FaceState{};
}

template <class F>
int fill_face_state(FaceState& fs, int nz, hipStream_t st) {
  const size_t ns = 2 * (size_t)nz + 2;
  k_fill_faces<F><<<blocks_for(fs.nup * ns, 256), 256, 0, st>>>(fs.nup * ns, (FaceRecF<F>*)fs.u,
                                                                 fs.nu, fs.mono_u);
  k_fill_faces<F><<<blocks_for(fs.nvp * ns, 256), 256, 0, st>>>(fs.nvp * ns, (FaceRecF<F>*)fs.v,
                                                                 fs.nv, fs.mono_v);
  HIP_TRY(hipGetLastError());
  fs.filled = true;
  return REDI_OK;
}

int fill_face_state_form(FaceState& fs, int nz, int form, hipStream_t st) {
  return form == REDI_FORM_FP32_STORE ? fill_face_state<Fp32>(fs, nz, st)
                                      : fill_face_state<Fp64>(fs, nz, st);
}

struct Tiles {
  int tx = 64, ty = 1;       // one 64-lane wavefront per tile
  int recon_layer_major = 0; // reconstruction layer-major (1) or per column (0)
  int group = kTracerGroup;  // tracers per flux launch (1..4)
  int recon_kt = kReconChunkDefault;  // layers per tiled-reconstruction workgroup
  int seg = 0;  // cells per run of the flux kernels (0: by domain size, flux_segment)
  int vertical = kTileVerticalDefault;  // sweep tiles numbered down rows first
};

TileMap make_tiles(int n0, int n1, const Tiles& t, unsigned* nblocks) {
  TileMap m;
  m.n0 = n0; m.n1 = n1; m.tx = t.tx; m.ty = t.ty;
  m.tiles0 = (n0 + t.tx - 1) / t.tx;
  m.tiles1 = (n1 + t.ty - 1) / t.ty;
  m.vertical = t.vertical;
  *nblocks = (unsigned)m.tiles0 * (unsigned)m.tiles1;
  return m;
}

// Phase A over both face grids (redi_calc_coeffs_x / _y).
// Opt-in slope taper of Phase A: 1 / face spacing of u and v faces and the
// slope limit (0 = off).
struct Taper {
  const double* inv_u = nullptr;
  const double* inv_v = nullptr;
  double slope_max = 0.0;
};

template <class F>
int launch_phase_a(int nx, int ny, int nz, const EosParams& eos, const typename F::R* h,
                   const typename F::R* t, const typename F::R* s, const double* wet_u,
                   const double* wet_v, const FaceState& fs, void* ws, const Tiles& tiles,
                   hipStream_t st, Timer& tm, const Taper& tp = Taper{}) {
  const int N2 = nx * ny;
  typename F::Col* rec = (typename F::Col*)ws;
  int slot = tm.begin(REDI_K_COLUMN_STATE, st);
  k_column_state<F><<<blocks_for(N2, 128), 128, 0, st>>>(ColumnArgs<F>{N2, nz, nx, h, t, s, eos, rec});
  tm.end(slot, st);
  const double pa_to_h = 1.0 / (kGravity * eos.rho0);
  const int bs = tiles.tx * tiles.ty;
  if (bs == 64 && nx > 1 && ny > 1) {  // both grids in one launch, one wave per grid
    SweepArgs<F> au{nx, nz, nx + 1, face_fa(2 * nz + 2), face_fb(2 * nz + 2), 1, 0, 1, 0, wet_u, rec,
                    (FaceRecF<F>*)fs.u, fs.mono_u, pa_to_h, tp.inv_u, tp.slope_max};
    SweepArgs<F> av{nx, nz, nx, face_fa(2 * nz + 2), face_fb(2 * nz + 2), 0, 1, 0, 1, wet_v, rec,
                    (FaceRecF<F>*)fs.v, fs.mono_v, pa_to_h, tp.inv_v, tp.slope_max};
    au.keep_inert = av.keep_inert = fs.filled ? 1 : 0;
    unsigned nbu, nbv;
    TileMap mu = make_tiles(nx - 1, ny, tiles, &nbu);
    TileMap mv = make_tiles(nx, ny - 1, tiles, &nbv);
    mu.tiles0 = mv.tiles0 = mu.tiles0 > mv.tiles0 ? mu.tiles0 : mv.tiles0;
    mu.tiles1 = mv.tiles1 = mu.tiles1 > mv.tiles1 ? mu.tiles1 : mv.tiles1;
    slot = tm.begin(REDI_K_SWEEP_UV, st);
    k_neutral_sweep_uv<F><<<(unsigned)mu.tiles0 * (unsigned)mu.tiles1, 128, 2 * kSweepStageBytes, st>>>(
        au, mu, av, mv);
    tm.end(slot, st);
    HIP_TRY(hipGetLastError());
    return REDI_OK;
  }
  if (nx > 1) {  // u faces i = 2..nx, j = 1..ny (1-based)
    SweepArgs<F> a{nx, nz, nx + 1, face_fa(2 * nz + 2), face_fb(2 * nz + 2), 1, 0, 1, 0, wet_u, rec,
                   (FaceRecF<F>*)fs.u, fs.mono_u, pa_to_h, tp.inv_u, tp.slope_max};
    a.keep_inert = fs.filled ? 1 : 0;
    unsigned nb;
    const TileMap m = make_tiles(nx - 1, ny, tiles, &nb);
    slot = tm.begin(REDI_K_SWEEP_U, st);
    k_neutral_sweep<F><<<nb, bs, (bs / 64) * kSweepStageBytes, st>>>(a, m);
    tm.end(slot, st);
  }
  if (ny > 1) {  // v faces j = 2..ny, i = 1..nx
    SweepArgs<F> a{nx, nz, nx, face_fa(2 * nz + 2), face_fb(2 * nz + 2), 0, 1, 0, 1, wet_v, rec,
                   (FaceRecF<F>*)fs.v, fs.mono_v, pa_to_h, tp.inv_v, tp.slope_max};
    a.keep_inert = fs.filled ? 1 : 0;
    unsigned nb;
    const TileMap m = make_tiles(nx, ny - 1, tiles, &nb);
    slot = tm.begin(REDI_K_SWEEP_V, st);
    k_neutral_sweep<F><<<nb, bs, (bs / 64) * kSweepStageBytes, st>>>(a, m);
    tm.end(slot, st);
  }
  HIP_TRY(hipGetLastError());
  return REDI_OK;
}

struct PhaseB {
  int nx, ny, nz, nghost, nxp, nyp;
  double dt;
  int ww, we, ws, wn;
  const double *khtr_u_ext, *khtr_v_ext;
  double khtr;
  const double *dy_cu, *dx_cv, *idxCu, *idyCv, *areaT;
};

template <class F, int NT>
void launch_recon(const PhaseB& p, const Tiles& tiles, const typename F::R* h,
                  const TracerSetT<typename F::R>& ts, typename F::Recon* recon, const ReconLayout& ry,
                  hipStream_t st) {
  const int N2 = p.nx * p.ny, nz = p.nz;
  const ReconArgs<F, NT> ra{N2, nz, h, ts, recon, ry};
  if (tiles.recon_layer_major) {
    k_tracer_recon<F, NT><<<blocks_for(N2, 128), 128, 0, st>>>(ra);
  } else {
    const unsigned nbr = recon_tiled_blocks(N2, nz, tiles.recon_kt);
    switch (tiles.recon_kt) {
      case 8: k_tracer_recon_tiled<F, NT, 8><<<nbr, kReconLanes, 0, st>>>(ra); break;
      case 32: k_tracer_recon_tiled<F, NT, 32><<<nbr, kReconLanes, 0, st>>>(ra); break;
      default: k_tracer_recon_tiled<F, NT, 16><<<nbr, kReconLanes, 0, st>>>(ra); break;
    }
  }
}

template <class F, int NT, int NZC>
void launch_flux_pair(const FluxArgs<F>& fx, const FluxArgs<F>& fy, unsigned nbx, unsigned nby,
                      hipStream_t st, Timer& tm) {
  int slot = tm.begin(REDI_K_FLUX_X, st);
  k_flux<F, NT, NZC, false><<<nbx, 64, 0, st>>>(fx);
  tm.end(slot, st);
  slot = tm.begin(REDI_K_FLUX_Y, st);
  k_flux<F, NT, NZC, true><<<nby, 64, 0, st>>>(fy);
  tm.end(slot, st);
}

// Smallest layer-capacity class of the flux kernels that holds nz layers.
int flux_class(int nz) {
  int c = 0;
  while (kFluxNzc[c] < nz) ++c;
  return c;
}

// One tracer group of Phase B: reconstruction, the fluxes of both face
// grids with their divergence, the tracer update.
template <class F, int NT>
void launch_group(const PhaseB& p, const Tiles& tiles, const typename F::R* h,
                  const TracerSetT<typename F::R>& ts, typename F::Recon* recon, const ReconLayout& ry,
                  const FluxArgs<F>& fx, const FluxArgs<F>& fy, unsigned nbx, unsigned nby,
                  const UpdateArgs<F>& ua, hipStream_t st, Timer& tm) {
  int slot = tm.begin(REDI_K_TRACER_RECON, st);
  launch_recon<F, NT>(p, tiles, h, ts, recon, ry, st);
  tm.end(slot, st);
  switch (flux_class(p.nz)) {
    case 0: launch_flux_pair<F, NT, kFluxNzc[0]>(fx, fy, nbx, nby, st, tm); break;
    case 1: launch_flux_pair<F, NT, kFluxNzc[1]>(fx, fy, nbx, nby, st, tm); break;
    case 2: launch_flux_pair<F, NT, kFluxNzc[2]>(fx, fy, nbx, nby, st, tm); break;
    default: launch_flux_pair<F, NT, kFluxNzc[3]>(fx, fy, nbx, nby, st, tm); break;
  }
  slot = tm.begin(REDI_K_TRACER_UPDATE, st);
  k_tracer_update<F><<<update_blocks(p.nx * p.ny, p.nz), kUpdLanes, 0, st>>>(ua);
  tm.end(slot, st);
}

// Cells per run of the flux kernels for an nx x ny domain: runs of
// kFluxSegDefault cells where that still gives every SIMD several runs to
// interleave, shorter runs (more of them, each end face computed twice) on
// small domains.
int flux_segment(int nx, int ny, int seg) {
  if (seg > 0) return seg;
  const double cells = (double)nx * ny;
  int s = kFluxSegDefault;
  while (s > 8 && cells / s < 20000.0) s /= 2;
  return s;
}

// Flux arguments of both face grids and the update arguments of one
// domain's Phase B over the scratch `sc` (the reconstruction layout, the
// tracer pointers and the group size are set per tracer group).
template <class F>
void phase_b_args(const PhaseB& p, const FaceState& fs, const PhaseBScratch& sc, int seg,
                  FluxArgs<F>& fx, FluxArgs<F>& fy, UpdateArgs<F>& ua, unsigned* nbx, unsigned* nby) {
  const int nx = p.nx, ny = p.ny, nz = p.nz, ns = 2 * nz + 2;
  typename F::Recon* recon = (typename F::Recon*)sc.recon;
  fx = FluxArgs<F>{};  // u faces: runs along rows
  fx.nx = nx; fx.ny = ny; fx.nz = nz;
  fx.wall0 = p.ww ? p.nghost : -1;
  fx.wall1 = p.we ? p.nghost + p.nxp : -1;
  seg = flux_segment(nx, ny, seg);
  fx.seg = seg;
  *nbx = flux_blocks(nx, ny, seg, false, &fx.nseg);
  fx.fa = face_fa(ns); fx.fb = face_fb(ns);
  fx.dt = p.dt; fx.khtr = fs.khtr_u; fx.len = p.dy_cu; fx.inv = p.idxCu;
  fx.rec = (const FaceRecF<F>*)fs.u; fx.flags = fs.mono_u;
  fx.recon = recon; fx.acc = sc.acc;
  fy = fx;  // v faces: runs up columns
  fy.wall0 = p.ws ? p.nghost : -1;
  fy.wall1 = p.wn ? p.nghost + p.nyp : -1;
  *nby = flux_blocks(nx, ny, seg, true, &fy.nseg);
  fy.khtr = fs.khtr_v; fy.len = p.dx_cv; fy.inv = p.idyCv;
  fy.rec = (const FaceRecF<F>*)fs.v; fy.flags = fs.mono_v;
  ua = UpdateArgs<F>{};
  ua.N2 = nx * ny; ua.nz = nz; ua.areaT = p.areaT; ua.acc = sc.acc;
}

// Reconstruction layout of a group of k tracers with ne entries per column
// and tracer (nz layers plus the form's recon_extra).
ReconLayout recon_layout(int layer_major, int N2, int ne, int k) {
  const uint32_t k32 = (uint32_t)k;
  return layer_major ? ReconLayout{k32, (uint32_t)N2 * k32, 1}
                     : ReconLayout{(uint32_t)ne * k32, 1, (uint32_t)ne};
}

// Sublayer chunks per face of the guard-decision output.
size_t guard_chunks(int nz) { return (size_t)(2 * nz + 1 + kChunk - 1) / kChunk; }

template <class F>
int launch_phase_b(const PhaseB& p, const typename F::R* h, int nt, typename F::R* const* tracers,
                   const FaceState& fs, void* ws, const Tiles& tiles, hipStream_t st, Timer& tm,
                   double* const* tend = nullptr, uint64_t* guard_u = nullptr,
                   uint64_t* guard_v = nullptr) {
  using R = typename F::R;
  const int nx = p.nx, ny = p.ny, nz = p.nz, N2 = nx * ny;
  const int nu = (nx + 1) * ny, nv = nx * (ny + 1);
  int slot = tm.begin(REDI_K_FACE_KHTR, st);
  k_face_khtr<<<blocks_for((size_t)nu + nv, 256), 256, 0, st>>>(
      KhtrArgs{nu, nv, p.khtr_u_ext, p.khtr_v_ext, p.khtr, fs.khtr_u, fs.khtr_v});
  tm.end(slot, st);
  const int slots = recon_slots(nt);
  const PhaseBScratch sc = carve_phase_b(ws, nx, ny, nz, slots, form_sizes<F>());
  typename F::Recon* recon = (typename F::Recon*)sc.recon;
  FluxArgs<F> fx, fy;
  UpdateArgs<F> ua;
  unsigned nbx = 0, nby = 0;
  phase_b_args<F>(p, fs, sc, tiles.seg, fx, fy, ua, &nbx, &nby);
  const int G = std::max(1, std::min(tiles.group, slots));
  for (int g0 = 0; g0 < nt; g0 += G) {
    const int gn = std::min(G, nt - g0);
    // A lone tracer goes through the two-tracer kernels: slot 1 repeats it
    // and its result lands in the discard field.
    const int kn = gn == 1 ? 2 : gn;
    TracerSetT<R> ts{};
    for (int t = 0; t < kMaxNT; ++t) {
      const int tt = std::min(t, gn - 1);
      ts.q[t] = tracers[g0 + tt];
      ua.htr[t] = tracers[g0 + tt];
      ua.tend[t] = tend ? tend[g0 + tt] : nullptr;
    }
    if (gn == 1) {
      ua.htr[1] = (R*)sc.discard;
      ua.tend[1] = nullptr;
    }
    ua.nt = kn;
    const ReconLayout ry = recon_layout(tiles.recon_layer_major, N2, nz + F::Recon::kExtra, kn);
    fx.ry = fy.ry = ry;
    if (guard_u && guard_v) {
      const size_t nch = guard_chunks(nz);
      fx.guard = guard_u + (size_t)g0 * nu * nch;
      fy.guard = guard_v + (size_t)g0 * nv * nch;
      fx.guard_ts = (size_t)nu * nch;
      fy.guard_ts = (size_t)nv * nch;
      fx.guard_fs = fy.guard_fs = nch;
      fx.guard_nt = fy.guard_nt = gn;
    }
    switch (kn) {
      case 2: launch_group<F, 2>(p, tiles, h, ts, recon, ry, fx, fy, nbx, nby, ua, st, tm); break;
      case 3: launch_group<F, 3>(p, tiles, h, ts, recon, ry, fx, fy, nbx, nby, ua, st, tm); break;
      default: launch_group<F, 4>(p, tiles, h, ts, recon, ry, fx, fy, nbx, nby, ua, st, tm); break;
    }
  }
  HIP_TRY(hipGetLastError());
  return REDI_OK;
}

// Write the reference-layout coefficient arrays (and khtr) into `d`.
// all_faces: every face (the reference's init values where Phase A never
// writes); otherwise only the faces redi_calc_coeffs_x/_y compute.
template <class F>
int export_coeffs(const FaceState& fs, int nx, int ny, int nz, const redi_coeffs_t& d,
                  bool all_faces, bool with_khtr, hipStream_t st) {
  const int ns = 2 * nz + 2;
  const size_t nu = (size_t)(nx + 1) * ny, nv = (size_t)nx * (ny + 1);
  const uint32_t ua = face_fa(ns), ub = face_fb(ns), va = face_fa(ns), vb = face_fb(ns);
  const FaceRecF<F>* ru = (const FaceRecF<F>*)fs.u;
  const FaceRecF<F>* rv = (const FaceRecF<F>*)fs.v;
  CoeffArrays cu{d.uPoL, d.uPoR, d.uhEff, d.uKoL, d.uKoR};
  CoeffArrays cv{d.vPoL, d.vPoR, d.vhEff, d.vKoL, d.vKoR};
  if (all_faces) {
    k_export_coeffs<F><<<blocks_for(nu * ns, 256), 256, 0, st>>>(ru, ua, ub, nx + 1, ny, ns, 0, nx + 1, 0, ny, cu);
    k_export_coeffs<F><<<blocks_for(nv * ns, 256), 256, 0, st>>>(rv, va, vb, nx, ny + 1, ns, 0, nx, 0, ny + 1, cv);
  } else {
    k_export_coeffs<F><<<blocks_for(nu * ns, 256), 256, 0, st>>>(ru, ua, ub, nx + 1, ny, ns, 1, nx, 0, ny, cu);
    k_export_coeffs<F><<<blocks_for(nv * ns, 256), 256, 0, st>>>(rv, va, vb, nx, ny + 1, ns, 0, nx, 1, ny, cv);
  }
  if (with_khtr) {
    if (d.khtr_u) HIP_TRY(hipMemcpyAsync(d.khtr_u, fs.khtr_u, nu * 8, hipMemcpyDeviceToDevice, st));
    if (d.khtr_v) HIP_TRY(hipMemcpyAsync(d.khtr_v, fs.khtr_v, nv * 8, hipMemcpyDeviceToDevice, st));
  }
  HIP_TRY(hipGetLastError());
  return REDI_OK;
}

EosParams to_eos(const redi_eos_t& e) {
  return EosParams{e.variant, e.ts_convention, e.rho0, e.T_ref, e.S_ref, e.alpha_T, e.beta_S, e.p_ref};
}

bool valid_eos(int v) {
  return v == REDI_EOS_LINEAR || v == REDI_EOS_WRIGHT97 || v == REDI_EOS_ROQUET_SPV;
}

// Per-device grow-only state for the flat entries (which, like the
// reference launchers, receive no state object).  FP64 form.
struct FlatSlot {
  int nx = 0, ny = 0, nz = 0;
  FaceState fs;
  void* ws = nullptr;
  size_t ws_bytes = 0;
};
std::mutex g_flat_mu;
std::vector<FlatSlot> g_flat;

int flat_state(int nx, int ny, int nz, int nt, FlatSlot** out) {
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(g_flat_mu);
  if ((int)g_flat.size() <= dev) g_flat.resize(dev + 1);
  FlatSlot& s = g_flat[dev];
  if (s.nx != nx || s.ny != ny || s.nz != nz) {
    if (s.fs.u) free_face_state(&s.fs, false);
    s.nx = s.ny = s.nz = 0;
    int rc = alloc_face_state(&s.fs, nx, ny, nz, false, REDI_FORM_FP64);
    if (rc) return rc;
    rc = fill_face_state<Fp64>(s.fs, nz, nullptr);
    if (rc) return rc;
    HIP_TRY(hipDeviceSynchronize());
    s.nx = nx; s.ny = ny; s.nz = nz;
  }
  const size_t wb = workspace_bytes(nx, ny, nz, nt, REDI_FORM_FP64);
  if (s.ws_bytes < wb) {
    if (s.ws) HIP_TRY(hipFree(s.ws));
    s.ws = nullptr; s.ws_bytes = 0;
    HIP_TRY(hipMalloc(&s.ws, wb));
    s.ws_bytes = wb;
  }
  *out = &s;
  return REDI_OK;
}


int check_flat(const char* who, int nx, int ny, int nz, const redi_eos_t* eos, int nt,
               double* const* tracers, int idx_T, int idx_S, const redi_coeffs_t* cf) {
  if (!eos || !cf || !tracers || nt < 1 || nx < 1 || ny < 1 || nz < 1)
    return fail(REDI_ERR_ARG, "%s: bad argument", who);
  if (nz > REDI_NZ_MAX) return fail(REDI_ERR_NZ, "%s: nz %d > %d", who, nz, REDI_NZ_MAX);
  if ((double)(nz + 1) * nx * ny * kTracerGroup >= 2147483648.0)
    return fail(REDI_ERR_ARG, "%s: domain exceeds 2^29 cell interfaces", who);
  if (idx_T < 1 || idx_T > nt || idx_S < 1 || idx_S > nt)
    return fail(REDI_ERR_ARG, "%s: tracer index out of range", who);
  if (!valid_eos(eos->variant)) return fail(REDI_ERR_ARG, "%s: unknown EOS variant", who);
  if (!cf->khtr_u || !cf->khtr_v) return fail(REDI_ERR_ARG, "%s: null khtr_u/khtr_v", who);
  return REDI_OK;
}

// One flat Redi step (FP64) over caller-owned fields and the face state /
// workspace `fs`, `ws`: Phase A, the reference-layout coefficients of the
// computed faces, Phase B.  The caller's khtr arrays take the resolved face
// diffusivities.
int launch_flat(int nx, int ny, int nz, double dt, int nghost, int nxp, int nyp, int ww, int we,
                int wsz, int wn, const redi_eos_t& eos, const double* h_layer, int nt,
                double* const* tracers, int idx_T, int idx_S, const double* wet_u,
                const double* wet_v, const redi_coeffs_t& cf, const double* kue, const double* kve,
                const double* dy_cu, const double* dx_cv, const double* idxCu, const double* idyCv,
                const double* areaT, FaceState fs, void* ws, hipStream_t st) {
  const Range range("redi_launch (flat)");
  fs.khtr_u = cf.khtr_u;
  fs.khtr_v = cf.khtr_v;
  const Tiles tiles;
  int rc = launch_phase_a<Fp64>(nx, ny, nz, to_eos(eos), h_layer, tracers[idx_T - 1],
                                tracers[idx_S - 1], wet_u, wet_v, fs, ws, tiles, st, g_no_timer);
  if (rc) return rc;
  rc = export_coeffs<Fp64>(fs, nx, ny, nz, cf, false, false, st);
  if (rc) return rc;
  PhaseB p;
  p.nx = nx; p.ny = ny; p.nz = nz; p.nghost = nghost; p.nxp = nxp; p.nyp = nyp; p.dt = dt;
  p.ww = ww; p.we = we; p.ws = wsz; p.wn = wn;
  p.khtr_u_ext = kue; p.khtr_v_ext = kve; p.khtr = 0.0;
  p.dy_cu = dy_cu; p.dx_cv = dx_cv; p.idxCu = idxCu; p.idyCv = idyCv; p.areaT = areaT;
  return launch_phase_b<Fp64>(p, h_layer, nt, tracers, fs, ws, tiles, st, g_no_timer);
}

// The allocation-free flat entry's scratch: one static array per device,
// holding the FP64 face records and flags of both face grids and the step
// workspace.  Its host-side name is not a device address; each device's
// address comes from hipGetSymbolAddress, once, with that device current.
// Every value a call reads from it, the call first writes (the sweep writes
// every computed face, Phase B every cell), so nothing in it needs clearing
// between calls; calls that overlap on one device race on it.
constexpr size_t kNoAllocScratchBytes = REDI_NOALLOC_SCRATCH_BYTES;
__device__ __align__(256) char scratch[kNoAllocScratchBytes];
std::mutex g_scratch_mu;
std::vector<char*> g_scratch_ptr;  // per device

size_t noalloc_bytes(int nx, int ny, int nz, int nt) {
  const size_t ns = 2 * (size_t)nz + 2, nu = (size_t)(nx + 1) * ny, nv = (size_t)nx * (ny + 1);
  auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
  return up(padded_faces(nu) * ns * sizeof(FaceRec64)) + up(padded_faces(nv) * ns * sizeof(FaceRec64)) +
         up(nu) + up(nv) + up(workspace_bytes(nx, ny, nz, nt, REDI_FORM_FP64));
}

int noalloc_scratch(char** out) {
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if ((int)g_scratch_ptr.size() <= dev) g_scratch_ptr.resize(dev + 1, nullptr);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return fail(REDI_ERR_HIP, "redi_launch_nt_noalloc: no device address for the scratch");
    g_scratch_ptr[dev] = (char*)p;
  }
  *out = g_scratch_ptr[dev];
  return REDI_OK;
}

}  // namespace

struct redi_s {
  int device = 0;
  int form = REDI_FORM_FP64;
  int nx = 0, ny = 0, nz = 0, ns = 0;
  int enable = 0, continuous = 1;
  double khtr = 0.0, ref_pres = -1.0;
  FaceState fs;
  void* ws = nullptr;
  size_t ws_bytes = 0;
  Tiles tiles;
  Timer timer;
  std::vector<double*> tend;  // optional tendency outputs, one per tracer
  uint64_t* guard_u = nullptr;  // optional guard-decision outputs (redi_set_guard_output)
  uint64_t* guard_v = nullptr;
  double slope_max = 0.0;     // opt-in slope taper (0 = off)
  bool is_init = false;
};

namespace {

// The model state of either form, reduced to what the launches need.
template <class R>
struct StateView {
  int nz_ml, idx_t, idx_s, nt;
  const R* h;
  R* const* tracers;
};

template <class F>
int calc_coeffs_impl(const char* who, const redi_hgrid_t* grid, const redi_metrics_t* m,
                     const redi_eos_t* eos, redi_t* r, const StateView<typename F::R>& ms,
                     void* stream) {
  const Range range(who);
  // silent guards (:585-593)
  if (!r || !r->is_init || !r->enable) return REDI_SKIPPED;
  if (ms.idx_t <= 0 || ms.idx_s <= 0) return REDI_SKIPPED;
  if (!ms.h) return REDI_SKIPPED;
  if (r->nz != ms.nz_ml) return REDI_SKIPPED;
  if (grid->nx_total != r->nx || grid->ny_total != r->ny)
    return fail(REDI_ERR_ARG, "%s: grid %dx%d does not match the state %dx%d", who,
                grid->nx_total, grid->ny_total, r->nx, r->ny);
  if (ms.idx_t > ms.nt || ms.idx_s > ms.nt || !ms.tracers)
    return fail(REDI_ERR_ARG, "%s: tracer index out of range", who);
  if (!valid_eos(eos->variant)) return fail(REDI_ERR_ARG, "%s: unknown EOS variant %d", who, eos->variant);
  if (!m->wet_u || !m->wet_v) return fail(REDI_ERR_ARG, "%s: null wet mask", who);
  Taper tp;
  if (r->slope_max > 0.0) {
    if (!m->idxCu || !m->idyCv) return fail(REDI_ERR_ARG, "%s: the slope taper needs idxCu and idyCv", who);
    tp = Taper{m->idxCu, m->idyCv, r->slope_max};
  }
  return launch_phase_a<F>(r->nx, r->ny, r->nz, to_eos(*eos), ms.h, ms.tracers[ms.idx_t - 1],
                           ms.tracers[ms.idx_s - 1], m->wet_u, m->wet_v, r->fs, r->ws, r->tiles,
                           (hipStream_t)stream, r->timer, tp);
}

template <class F>
int apply_flux_impl(const char* who, const redi_hgrid_t* grid, const redi_metrics_t* m, redi_t* r,
                    const StateView<typename F::R>& ms, double dt, const double* kue,
                    const double* kve, const redi_bc_t* bc, void* stream) {
  const Range range(who);
  // silent guards (:951-959, :978)
  if (!r || !r->is_init || !r->enable) return REDI_SKIPPED;
  if (!ms.h || !ms.tracers || ms.nt < 1) return REDI_SKIPPED;
  if (r->nz != ms.nz_ml) return REDI_SKIPPED;
  const bool use_ext = kue && kve;
  if (!use_ext && r->khtr <= 0.0) return REDI_SKIPPED;
  if (grid->nx_total != r->nx || grid->ny_total != r->ny)
    return fail(REDI_ERR_ARG, "%s: grid %dx%d does not match the state %dx%d", who,
                grid->nx_total, grid->ny_total, r->nx, r->ny);
  if (!m->dy_cu || !m->dx_cv || !m->idxCu || !m->idyCv || !m->areaT)
    return fail(REDI_ERR_ARG, "%s: null metric", who);
  for (int t = 0; t < ms.nt; ++t)
    if (!ms.tracers[t]) return fail(REDI_ERR_ARG, "%s: tracer %d is null", who, t + 1);
  PhaseB p;
  p.nx = r->nx; p.ny = r->ny; p.nz = r->nz;
  p.nghost = grid->nghost; p.nxp = grid->nx_phys; p.nyp = grid->ny_phys;
  p.dt = dt;
  p.ww = bc ? (bc->west == REDI_OBC_WALL) : 1;
  p.we = bc ? (bc->east == REDI_OBC_WALL) : 1;
  p.ws = bc ? (bc->south == REDI_OBC_WALL) : 1;
  p.wn = bc ? (bc->north == REDI_OBC_WALL) : 1;
  p.khtr_u_ext = use_ext ? kue : nullptr;
  p.khtr_v_ext = use_ext ? kve : nullptr;
  p.khtr = r->khtr;
  p.dy_cu = m->dy_cu; p.dx_cv = m->dx_cv; p.idxCu = m->idxCu; p.idyCv = m->idyCv;
  p.areaT = m->areaT;
  if (!r->tend.empty() && (int)r->tend.size() != ms.nt)
    return fail(REDI_ERR_ARG, "%s: %zu tendency outputs for %d tracers", who, r->tend.size(), ms.nt);
  return launch_phase_b<F>(p, ms.h, ms.nt, ms.tracers, r->fs, r->ws, r->tiles,
                           (hipStream_t)stream, r->timer, r->tend.empty() ? nullptr : r->tend.data(),
                           r->guard_u, r->guard_v);
}

StateView<double> view64(const redi_state_t& s) {
  return StateView<double>{s.nz_ml, s.idx_temperature, s.idx_salinity, s.ntracers, s.h_layer,
                           s.tracers};
}
StateView<float> view32(const redi_state32_t& s) {
  return StateView<float>{s.nz_ml, s.idx_temperature, s.idx_salinity, s.ntracers, s.h_layer,
                          s.tracers};
}

int wrong_form(const char* who, const redi_t* r, int want) {
  return fail(REDI_ERR_ARG, "%s: the state was created for form %d, not %d", who, r->form, want);
}

}  // namespace

extern "C" {

const char* redi_last_error(void) { return g_err.c_str(); }

int redi_version(void) { return 500; }

long long redi_workspace_bytes(int nx, int ny, int nz) {
  if (nx < 1 || ny < 1 || nz < 1) return 0;
  return (long long)workspace_bytes(nx, ny, nz, kTracerGroup, REDI_FORM_FP64);
}

long long redi_state_bytes(int nx, int ny, int nz) {
  return redi_state_bytes_form(nx, ny, nz, REDI_FORM_FP64);
}

long long redi_state_bytes_form(int nx, int ny, int nz, int form) {
  if (nx < 1 || ny < 1 || nz < 1) return 0;
  if (form != REDI_FORM_FP64 && form != REDI_FORM_FP32_STORE) return 0;
  return (long long)(face_state_bytes(nx, ny, nz, form) +
                     workspace_bytes(nx, ny, nz, kTracerGroup, form));
}

int redi_init(redi_t** out, const redi_hgrid_t* grid, int nz_ml) {
  return redi_init_form(out, grid, nz_ml, REDI_FORM_FP64);
}

int redi_init_form(redi_t** out, const redi_hgrid_t* grid, int nz_ml, int form) {
  if (!out || !grid) return fail(REDI_ERR_ARG, "redi_init: null argument");
  *out = nullptr;
  if (form != REDI_FORM_FP64 && form != REDI_FORM_FP32_STORE)
    return fail(REDI_ERR_ARG, "redi_init: unknown form %d", form);
  const int nx = grid->nx_total, ny = grid->ny_total;
  if (nx < 1 || ny < 1) return fail(REDI_ERR_ARG, "redi_init: bad grid %d x %d", nx, ny);
  const int nz = nz_ml < 1 ? 1 : nz_ml;
  if (nz > REDI_NZ_MAX)
    return fail(REDI_ERR_NZ, "redi_init: 2*nz_ml+2 = %d exceeds the neutral-surface limit %d",
                2 * nz + 2, 2 * REDI_NZ_MAX + 2);
  if ((double)(nz + 1) * nx * ny * kTracerGroup >= 2147483648.0)
    return fail(REDI_ERR_ARG, "redi_init: %d x %d x %d exceeds 2^29 cell interfaces per domain",
                nx, ny, nz);
  redi_t* r = new redi_t();
  r->form = form;
  auto cleanup = [&](int code) {
    free_face_state(&r->fs, true);
    (void)hipFree(r->ws);
    delete r;
    return code;
  };
  if (hipGetDevice(&r->device) != hipSuccess) return cleanup(fail(REDI_ERR_HIP, "redi_init: no device"));
  r->nx = nx; r->ny = ny; r->nz = nz; r->ns = 2 * nz + 2;
  int rc = alloc_face_state(&r->fs, nx, ny, nz, true, form);
  if (rc) return cleanup(rc);
  r->ws_bytes = workspace_bytes(nx, ny, nz, kTracerGroup, form);
  if (hipMalloc(&r->ws, r->ws_bytes) != hipSuccess)
    return cleanup(fail(REDI_ERR_HIP, "redi_init: workspace hipMalloc(%zu B) failed", r->ws_bytes));
  rc = fill_face_state_form(r->fs, nz, form, nullptr);
  if (rc) return cleanup(rc);
  if (hipMemset(r->fs.khtr_u, 0, r->fs.nu * 8) != hipSuccess ||
      hipMemset(r->fs.khtr_v, 0, r->fs.nv * 8) != hipSuccess || hipDeviceSynchronize() != hipSuccess)
    return cleanup(fail(REDI_ERR_HIP, "redi_init: initial fill failed"));
  r->is_init = true;
  *out = r;
  return REDI_OK;
}

int redi_destroy(redi_t* r) {
  if (!r) return REDI_OK;
  (void)hipDeviceSynchronize();
  free_face_state(&r->fs, true);
  (void)hipFree(r->ws);
  delete r;
  return REDI_OK;
}

int redi_configure(redi_t* r, int enable, double khtr, int continuous, double ref_pres) {
  if (!r) return fail(REDI_ERR_ARG, "redi_configure: null handle");
  if (!continuous)
    return fail(REDI_ERR_CONFIG, "redi_configure: only the continuous neutral-diffusion variant exists");
  r->enable = enable ? 1 : 0;
  r->khtr = khtr;
  r->continuous = 1;
  r->ref_pres = ref_pres;
  return REDI_OK;
}

int redi_set_tiles(redi_t* r, int tx, int ty) {
  if (!r) return fail(REDI_ERR_ARG, "redi_set_tiles: null handle");
  if (tx < 1 || ty < 1 || tx * ty > 256 || (tx * ty) % 64 != 0)
    return fail(REDI_ERR_ARG, "redi_set_tiles: tile %dx%d must hold 64, 128, 192 or 256 lanes", tx, ty);
  r->tiles.tx = tx;
  r->tiles.ty = ty;
  return REDI_OK;
}

int redi_set_tracer_group(redi_t* r, int n) {
  if (!r) return fail(REDI_ERR_ARG, "redi_set_tracer_group: null handle");
  if (n < 1 || n > kTracerGroup)
    return fail(REDI_ERR_ARG, "redi_set_tracer_group: %d not in 1..%d", n, kTracerGroup);
  r->tiles.group = n;
  return REDI_OK;
}

int redi_set_tendency_output(redi_t* r, int ntracers, double* const* tend) {
  if (!r) return fail(REDI_ERR_ARG, "redi_set_tendency_output: null handle");
  if (ntracers < 0 || (ntracers > 0 && !tend))
    return fail(REDI_ERR_ARG, "redi_set_tendency_output: bad arguments");
  r->tend.assign(tend, tend + ntracers);
  return REDI_OK;
}

int redi_set_guard_output(redi_t* r, uint64_t* u_bits, uint64_t* v_bits) {
  if (!r) return fail(REDI_ERR_ARG, "redi_set_guard_output: null handle");
  if (!u_bits != !v_bits) return fail(REDI_ERR_ARG, "redi_set_guard_output: give both arrays or neither");
  r->guard_u = u_bits;
  r->guard_v = v_bits;
  return REDI_OK;
}

int redi_set_slope_taper(redi_t* r, double slope_max) {
  if (!r) return fail(REDI_ERR_ARG, "redi_set_slope_taper: null handle");
  if (!(slope_max >= 0.0)) return fail(REDI_ERR_ARG, "redi_set_slope_taper: slope limit %g < 0", slope_max);
  r->slope_max = slope_max;
  return REDI_OK;
}

int redi_set_recon_chunk(redi_t* r, int kt) {
  if (!r) return fail(REDI_ERR_ARG, "redi_set_recon_chunk: null handle");
  if (kt != 8 && kt != 16 && kt != 32)
    return fail(REDI_ERR_ARG, "redi_set_recon_chunk: layer chunk %d is not 8, 16 or 32", kt);
  r->tiles.recon_kt = kt;
  return REDI_OK;
}

int redi_set_tile_order(redi_t* r, int vertical) {
  if (!r) return fail(REDI_ERR_ARG, "redi_set_tile_order: null handle");
  r->tiles.vertical = vertical ? 1 : 0;
  return REDI_OK;
}

int redi_set_flux_segment(redi_t* r, int cells) {
  if (!r) return fail(REDI_ERR_ARG, "redi_set_flux_segment: null handle");
  if (cells < 0 || cells > 4096)
    return fail(REDI_ERR_ARG, "redi_set_flux_segment: %d cells per run not in 0..4096", cells);
  r->tiles.seg = cells;
  return REDI_OK;
}

int redi_set_recon_order(redi_t* r, int layer_major) {
  if (!r) return fail(REDI_ERR_ARG, "redi_set_recon_order: null handle");
  r->tiles.recon_layer_major = layer_major ? 1 : 0;
  return REDI_OK;
}

int redi_nonmonotone_faces(const redi_t* r, long long* nu, long long* nv) {
  if (!r || !nu || !nv) return fail(REDI_ERR_ARG, "redi_nonmonotone_faces: null argument");
  std::vector<uint8_t> hu(r->fs.nu), hv(r->fs.nv);
  HIP_TRY(hipDeviceSynchronize());
  HIP_TRY(hipMemcpy(hu.data(), r->fs.mono_u, r->fs.nu, hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(hv.data(), r->fs.mono_v, r->fs.nv, hipMemcpyDeviceToHost));
  *nu = 0; *nv = 0;
  for (uint8_t f : hu) *nu += ((f & (kMonoL | kMonoR)) != (kMonoL | kMonoR));
  for (uint8_t f : hv) *nv += ((f & (kMonoL | kMonoR)) != (kMonoL | kMonoR));
  return REDI_OK;
}

int redi_export_coeffs(const redi_t* r, const redi_coeffs_t* dst, void* stream) {
  if (!r || !dst) return fail(REDI_ERR_ARG, "redi_export_coeffs: null argument");
  if (!dst->uPoL || !dst->uPoR || !dst->uKoL || !dst->uKoR || !dst->uhEff || !dst->vPoL ||
      !dst->vPoR || !dst->vKoL || !dst->vKoR || !dst->vhEff)
    return fail(REDI_ERR_ARG, "redi_export_coeffs: incomplete destination arrays");
  if (r->form == REDI_FORM_FP32_STORE)
    return export_coeffs<Fp32>(r->fs, r->nx, r->ny, r->nz, *dst, true, true, (hipStream_t)stream);
  return export_coeffs<Fp64>(r->fs, r->nx, r->ny, r->nz, *dst, true, true, (hipStream_t)stream);
}

int redi_calc_coeffs(const redi_hgrid_t* grid, const redi_metrics_t* m, const redi_eos_t* eos,
                     redi_t* r, const redi_state_t* ms, void* stream) {
  if (!grid || !m || !eos || !ms) return fail(REDI_ERR_ARG, "redi_calc_coeffs: null argument");
  if (r && r->form != REDI_FORM_FP64) return wrong_form("redi_calc_coeffs", r, REDI_FORM_FP64);
  return calc_coeffs_impl<Fp64>("redi_calc_coeffs", grid, m, eos, r, view64(*ms), stream);
}

int redi_calc_coeffs_fp32(const redi_hgrid_t* grid, const redi_metrics_t* m,
                          const redi_eos_t* eos, redi_t* r, const redi_state32_t* ms,
                          void* stream) {
  if (!grid || !m || !eos || !ms) return fail(REDI_ERR_ARG, "redi_calc_coeffs_fp32: null argument");
  if (r && r->form != REDI_FORM_FP32_STORE)
    return wrong_form("redi_calc_coeffs_fp32", r, REDI_FORM_FP32_STORE);
  return calc_coeffs_impl<Fp32>("redi_calc_coeffs_fp32", grid, m, eos, r, view32(*ms), stream);
}

int redi_apply_flux(const redi_hgrid_t* grid, const redi_metrics_t* m, redi_t* r,
                    const redi_state_t* ms, double dt, const double* kue, const double* kve,
                    const redi_bc_t* bc, void* stream) {
  if (!grid || !m || !ms) return fail(REDI_ERR_ARG, "redi_apply_flux: null argument");
  if (r && r->form != REDI_FORM_FP64) return wrong_form("redi_apply_flux", r, REDI_FORM_FP64);
  return apply_flux_impl<Fp64>("redi_apply_flux", grid, m, r, view64(*ms), dt, kue, kve, bc, stream);
}

int redi_apply_flux_fp32(const redi_hgrid_t* grid, const redi_metrics_t* m, redi_t* r,
                         const redi_state32_t* ms, double dt, const double* kue,
                         const double* kve, const redi_bc_t* bc, void* stream) {
  if (!grid || !m || !ms) return fail(REDI_ERR_ARG, "redi_apply_flux_fp32: null argument");
  if (r && r->form != REDI_FORM_FP32_STORE)
    return wrong_form("redi_apply_flux_fp32", r, REDI_FORM_FP32_STORE);
  return apply_flux_impl<Fp32>("redi_apply_flux_fp32", grid, m, r, view32(*ms), dt, kue, kve, bc,
                               stream);
}

int redi_apply_flux_hoist(const redi_hgrid_t* grid, const redi_metrics_t* m, redi_t* r,
                          const redi_state_t* ms, double dt, const double* kue, const double* kve,
                          const redi_bc_t* bc, void* stream) {
  return redi_apply_flux(grid, m, r, ms, dt, kue, kve, bc, stream);
}

int redi_step(const redi_hgrid_t* grid, const redi_metrics_t* m, const redi_eos_t* eos, redi_t* r,
              const redi_state_t* ms, double dt, const double* kue, const double* kve,
              const redi_bc_t* bc, int nsteps, void* stream) {
  int rc = REDI_OK;
  for (int n = 0; n < nsteps; ++n) {
    const int a = redi_calc_coeffs(grid, m, eos, r, ms, stream);
    if (a < 0) return a;
    const int b = redi_apply_flux(grid, m, r, ms, dt, kue, kve, bc, stream);
    if (b < 0) return b;
    if (a == REDI_SKIPPED && b == REDI_SKIPPED) rc = REDI_SKIPPED;
  }
  return rc;
}

int redi_step_fp32(const redi_hgrid_t* grid, const redi_metrics_t* m, const redi_eos_t* eos,
                   redi_t* r, const redi_state32_t* ms, double dt, const double* kue,
                   const double* kve, const redi_bc_t* bc, int nsteps, void* stream) {
  int rc = REDI_OK;
  for (int n = 0; n < nsteps; ++n) {
    const int a = redi_calc_coeffs_fp32(grid, m, eos, r, ms, stream);
    if (a < 0) return a;
    const int b = redi_apply_flux_fp32(grid, m, r, ms, dt, kue, kve, bc, stream);
    if (b < 0) return b;
    if (a == REDI_SKIPPED && b == REDI_SKIPPED) rc = REDI_SKIPPED;
  }
  return rc;
}

int redi_launch_nt(int nx, int ny, int nz, double dt, int nghost, int nxp, int nyp, int ww, int we,
                   int wsz, int wn, const redi_eos_t* eos, const double* h_layer, int nt,
                   double* const* tracers, int idx_T, int idx_S, const double* wet_u,
                   const double* wet_v, const redi_coeffs_t* cf, const double* kue,
                   const double* kve, const double* dy_cu, const double* dx_cv, const double* idxCu,
                   const double* idyCv, const double* areaT, void* stream) {
  int rc = check_flat("redi_launch_nt", nx, ny, nz, eos, nt, tracers, idx_T, idx_S, cf);
  if (rc) return rc;
  FlatSlot* fsl = nullptr;
  rc = flat_state(nx, ny, nz, nt, &fsl);
  if (rc) return rc;
  return launch_flat(nx, ny, nz, dt, nghost, nxp, nyp, ww, we, wsz, wn, *eos, h_layer, nt, tracers,
                     idx_T, idx_S, wet_u, wet_v, *cf, kue, kve, dy_cu, dx_cv, idxCu, idyCv, areaT,
                     fsl->fs, fsl->ws, (hipStream_t)stream);
}

int redi_launch_nt_noalloc(int nx, int ny, int nz, double dt, int nghost, int nxp, int nyp, int ww,
                           int we, int wsz, int wn, const redi_eos_t* eos, const double* h_layer,
                           int nt, double* const* tracers, int idx_T, int idx_S,
                           const double* wet_u, const double* wet_v, const redi_coeffs_t* cf,
                           const double* kue, const double* kve, const double* dy_cu,
                           const double* dx_cv, const double* idxCu, const double* idyCv,
                           const double* areaT, void* stream) {
  int rc = check_flat("redi_launch_nt_noalloc", nx, ny, nz, eos, nt, tracers, idx_T, idx_S, cf);
  if (rc) return rc;
  const size_t need = noalloc_bytes(nx, ny, nz, nt);
  if (need > kNoAllocScratchBytes)
    return fail(REDI_ERR_BUDGET,
                "redi_launch_nt_noalloc: %d x %d x %d with %d tracers needs %zu B of scratch; the "
                "static scratch holds %zu B",
                nx, ny, nz, nt, need, kNoAllocScratchBytes);
  char* base = nullptr;
  rc = noalloc_scratch(&base);
  if (rc) return rc;
  FaceState fs;
  const size_t ns = 2 * (size_t)nz + 2, rb = sizeof(FaceRec64);
  fs.nu = (size_t)(nx + 1) * ny;
  fs.nv = (size_t)nx * (ny + 1);
  fs.nup = padded_faces(fs.nu);
  fs.nvp = padded_faces(fs.nv);
  size_t off = 0;
  auto take = [&](size_t bytes) {
    char* p = base + off;
    off += (bytes + 255) & ~(size_t)255;
    return p;
  };
  fs.u = take(fs.nup * ns * rb);
  fs.v = take(fs.nvp * ns * rb);
  fs.mono_u = (uint8_t*)take(fs.nu);
  fs.mono_v = (uint8_t*)take(fs.nv);
  void* ws = take(workspace_bytes(nx, ny, nz, nt, REDI_FORM_FP64));
  return launch_flat(nx, ny, nz, dt, nghost, nxp, nyp, ww, we, wsz, wn, *eos, h_layer, nt, tracers,
                     idx_T, idx_S, wet_u, wet_v, *cf, kue, kve, dy_cu, dx_cv, idxCu, idyCv, areaT,
                     fs, ws, (hipStream_t)stream);
}

int redi_timing_enable(redi_t* r, int on) {
  if (!r) return fail(REDI_ERR_ARG, "redi_timing_enable: null handle");
  r->timer.on = on != 0;
  return REDI_OK;
}

int redi_timing_collect(redi_t* r, redi_kernel_times_t* out) {
  if (!r || !out) return fail(REDI_ERR_ARG, "redi_timing_collect: null argument");
  for (int k = 0; k < REDI_NKERNELS; ++k) { out->ms[k] = 0.0; out->launches[k] = 0; }
  for (auto& rec : r->timer.live) {
    HIP_TRY(hipEventSynchronize(rec.b));
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, rec.a, rec.b));
    out->ms[rec.kind] += ms;
    out->launches[rec.kind] += 1;
    r->timer.pool.push_back(rec.a);
    r->timer.pool.push_back(rec.b);
  }
  r->timer.live.clear();
  return REDI_OK;
}

const char* redi_kernel_name(int kind) {
  static const char* names[REDI_NKERNELS] = {"k_column_state", "k_neutral_sweep_u",
                                             "k_neutral_sweep_v", "k_face_khtr",
                                             "k_tracer_recon", "k_flux_x",
                                             "k_flux_y", "k_tracer_update",
                                             "k_neutral_sweep_uv"};
  return (kind >= 0 && kind < REDI_NKERNELS) ? names[kind] : "";
}

static void flat_launch(const int* nx, const int* ny, const int* nz, const double* dt,
                        const int* nghost, const int* nxp, const int* nyp, const int* ww,
                        const int* we, const int* wsz, const int* wn, const double* rho0,
                        const double* T_ref, const double* S_ref, const double* alpha_T,
                        const double* beta_S, const double* h_layer, double* t_htr, double* s_htr,
                        const double* wet_u, const double* wet_v, double* uPoL, double* uPoR,
                        int* uKoL, int* uKoR, double* uhEff, double* vPoL, double* vPoR, int* vKoL,
                        int* vKoR, double* vhEff, const double* khtr_u_ext,
                        const double* khtr_v_ext, double* khtr_u, double* khtr_v,
                        const double* dy_cu, const double* dx_cv, const double* idxCu,
                        const double* idyCv, const double* areaT) {
  redi_eos_t eos{REDI_EOS_LINEAR, 1, *rho0, *T_ref, *S_ref, *alpha_T, *beta_S, 0.0};
  redi_coeffs_t cf{uPoL, uPoR, uKoL, uKoR, uhEff, vPoL, vPoR, vKoL, vKoR, vhEff, khtr_u, khtr_v};
  double* tr[2] = {t_htr, s_htr};
  const int rc = redi_launch_nt(*nx, *ny, *nz, *dt, *nghost, *nxp, *nyp, *ww, *we, *wsz, *wn, &eos,
                                h_layer, 2, tr, 1, 2, wet_u, wet_v, &cf, khtr_u_ext, khtr_v_ext,
                                dy_cu, dx_cv, idxCu, idyCv, areaT, nullptr);
  if (rc < 0) fprintf(stderr, "redi flat launch: %s\n", g_err.c_str());
  (void)hipDeviceSynchronize();
}

void redi_opt_launch(const int* nx, const int* ny, const int* nz, const int* /*ns*/,
                     const double* dt, const int* nghost, const int* nxp, const int* nyp,
                     const int* ww, const int* we, const int* wsz, const int* wn,
                     const double* rho0, const double* T_ref, const double* S_ref,
                     const double* alpha_T, const double* beta_S, const double* h_layer,
                     double* t_htr, double* s_htr, const double* wet_u, const double* wet_v,
                     double* uPoL, double* uPoR, int* uKoL, int* uKoR, double* uhEff, double* vPoL,
                     double* vPoR, int* vKoL, int* vKoR, double* vhEff, const double* khtr_u_ext,
                     const double* khtr_v_ext, double* khtr_u, double* khtr_v,
                     const double* dy_cu, const double* dx_cv, const double* idxCu,
                     const double* idyCv, const double* areaT, double* /*tr_snap*/) {
  flat_launch(nx, ny, nz, dt, nghost, nxp, nyp, ww, we, wsz, wn, rho0, T_ref, S_ref, alpha_T,
              beta_S, h_layer, t_htr, s_htr, wet_u, wet_v, uPoL, uPoR, uKoL, uKoR, uhEff, vPoL,
              vPoR, vKoL, vKoR, vhEff, khtr_u_ext, khtr_v_ext, khtr_u, khtr_v, dy_cu, dx_cv, idxCu,
              idyCv, areaT);
}

void redi_cuda_launch_(const int* nx, const int* ny, const int* nz, const int* ns,
                       const double* dt, const int* nghost, const int* nxp, const int* nyp,
                       const int* ww, const int* we, const int* wsz, const int* wn,
                       const double* rho0, const double* T_ref, const double* S_ref,
                       const double* alpha_T, const double* beta_S, const double* h_layer,
                       double* t_htr, double* s_htr, const double* wet_u, const double* wet_v,
                       double* uPoL, double* uPoR, int* uKoL, int* uKoR, double* uhEff,
                       double* vPoL, double* vPoR, int* vKoL, int* vKoR, double* vhEff,
                       const double* khtr_u_ext, const double* khtr_v_ext, double* khtr_u,
                       double* khtr_v, const double* dy_cu, const double* dx_cv,
                       const double* idxCu, const double* idyCv, const double* areaT,
                       double* tr_snap) {
  redi_opt_launch(nx, ny, nz, ns, dt, nghost, nxp, nyp, ww, we, wsz, wn, rho0, T_ref, S_ref,
                  alpha_T, beta_S, h_layer, t_htr, s_htr, wet_u, wet_v, uPoL, uPoR, uKoL, uKoR,
                  uhEff, vPoL, vPoR, vKoL, vKoR, vhEff, khtr_u_ext, khtr_v_ext, khtr_u, khtr_v,
                  dy_cu, dx_cv, idxCu, idyCv, areaT, tr_snap);
}

}  // extern "C"

// The batched form (redi_batch_*).
#include "redi_batch_host.hpp"
