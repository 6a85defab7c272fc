// redi_lib.hip -- the C ABI of include/redi_hip.h over the gfx950 kernels.
//
// Host-side restatement of the reference's outer shims: ocean_redi_init /
// destroy (redi/ocean_redi.F90:466-529), redi_calc_coeffs (:572-607) and
// redi_apply_flux (:927-1002) with the same silent no-op guards, plus the
// flat launchers with the argument list of redi/opt_kernel.cu:534-574.
#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <mutex>
#include <string>
#include <vector>

#include "../../include/redi_hip.h"
#include "redi_kernels_fp64.hpp"

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
  // Records the start event; returns the slot to close with end().
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

// Column scratch of one step: Phase A holds 5 interface arrays, Phase B 4
// reconstruction arrays (they are live at different times and share it).
size_t workspace_elems(int nx, int ny, int nz) { return (size_t)5 * (nz + 1) * nx * ny; }

struct Workspace {
  double* base = nullptr;
  size_t elems = 0;
  ColumnState column(int N2, int nz) const {
    const size_t s = (size_t)(nz + 1) * N2;
    return ColumnState{base, base + s, base + 2 * s, base + 3 * s, base + 4 * s};
  }
  TracerRecon recon(int N2, int nz) const {
    const size_t s = (size_t)(nz + 1) * N2;
    return TracerRecon{base, base + s, base + 2 * s, base + 3 * s};
  }
};

// Phase A over both face grids (redi_calc_coeffs_x / _y).
int launch_phase_a(int nx, int ny, int nz, const EosParams& eos, const double* h, const double* t,
                   const double* s, const double* wet_u, const double* wet_v, const FaceCoeffs& u,
                   const FaceCoeffs& v, const Workspace& ws, hipStream_t st, Timer& tm) {
  const int N2 = nx * ny;
  const ColumnState cs = ws.column(N2, nz);
  int slot = tm.begin(REDI_K_COLUMN_STATE, st);
  k_column_state<<<blocks_for(N2, 128), 128, 0, st>>>(N2, nz, h, t, s, eos, cs);
  tm.end(slot, st);
  const double pa_to_h = 1.0 / (kGravity * eos.rho0);
  // u faces i = 2..nx, j = 1..ny (1-based) -> 0-based i = 1..nx-1
  if (nx > 1) {
    const long nf = (long)(nx - 1) * ny;
    slot = tm.begin(REDI_K_SWEEP_U, st);
    k_neutral_sweep<<<blocks_for(nf, 64), 64, 0, st>>>(nx, N2, nz, nx + 1, ny, 1, 0, nx - 1, ny, 1, 0,
                                                       wet_u, cs, u, pa_to_h);
    tm.end(slot, st);
  }
  // v faces j = 2..ny, i = 1..nx
  if (ny > 1) {
    const long nf = (long)nx * (ny - 1);
    slot = tm.begin(REDI_K_SWEEP_V, st);
    k_neutral_sweep<<<blocks_for(nf, 64), 64, 0, st>>>(nx, N2, nz, nx, ny + 1, 0, 1, nx, ny - 1, 0, 1,
                                                       wet_v, cs, v, pa_to_h);
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
  double *khtr_u, *khtr_v;
  const double *dy_cu, *dx_cv, *idxCu, *idyCv, *areaT, *h;
  FaceCoeffs u, v;
};

int flux_block(int nz) { return nz <= 128 ? 64 : 32; }

int launch_phase_b(const PhaseB& p, int nt, double* const* tracers, const Workspace& ws,
                   hipStream_t st, Timer& tm) {
  const int nx = p.nx, ny = p.ny, nz = p.nz, N2 = nx * ny;
  const int nu = (nx + 1) * ny, nv = nx * (ny + 1);
  int slot = tm.begin(REDI_K_FACE_KHTR, st);
  k_face_khtr<<<blocks_for((size_t)nu + nv, 256), 256, 0, st>>>(
      nu, nv, p.khtr_u_ext, p.khtr_v_ext, p.khtr, p.khtr_u, p.khtr_v);
  tm.end(slot, st);
  FluxArgs fa;
  fa.nx = nx; fa.ny = ny; fa.nz = nz; fa.dt = p.dt;
  fa.wuf_w = p.nghost + 1; fa.wuf_e = p.nghost + p.nxp + 1;
  fa.wvf_s = p.nghost + 1; fa.wvf_n = p.nghost + p.nyp + 1;
  fa.wall_w = p.ww; fa.wall_e = p.we; fa.wall_s = p.ws; fa.wall_n = p.wn;
  fa.khtr_u = p.khtr_u; fa.khtr_v = p.khtr_v;
  fa.dy_cu = p.dy_cu; fa.dx_cv = p.dx_cv; fa.idxCu = p.idxCu; fa.idyCv = p.idyCv;
  fa.areaT = p.areaT;
  fa.u = p.u; fa.v = p.v;
  const TracerRecon rc = ws.recon(N2, nz);
  const int bs = flux_block(nz);
  const size_t lds = (size_t)nz * bs * sizeof(double);
  for (int t = 0; t < nt; ++t) {
    slot = tm.begin(REDI_K_TRACER_RECON, st);
    k_tracer_recon<<<blocks_for(N2, 128), 128, 0, st>>>(N2, nz, p.h, tracers[t], rc);
    tm.end(slot, st);
    slot = tm.begin(REDI_K_APPLY_FLUX, st);
    k_apply_flux<<<blocks_for(N2, bs), bs, lds, st>>>(fa, rc, p.h, tracers[t]);
    tm.end(slot, st);
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

// Per-device grow-only scratch for the flat entries (which, like the
// reference launchers, receive no state object).
struct FlatCache {
  std::mutex mu;
  std::vector<std::pair<double*, size_t>> per_dev;
};
FlatCache g_flat;

int flat_workspace(size_t elems, Workspace* out) {
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(g_flat.mu);
  if ((int)g_flat.per_dev.size() <= dev) g_flat.per_dev.resize(dev + 1, {nullptr, 0});
  auto& slot = g_flat.per_dev[dev];
  if (slot.second < elems) {
    if (slot.first) HIP_TRY(hipFree(slot.first));
    slot = {nullptr, 0};
    HIP_TRY(hipMalloc(&slot.first, elems * sizeof(double)));
    slot.second = elems;
  }
  out->base = slot.first;
  out->elems = slot.second;
  return REDI_OK;
}

}  // namespace

struct redi_s {
  int device = 0;
  int nx = 0, ny = 0, nz = 0, ns = 0;
  int enable = 0, continuous = 1;
  double khtr = 0.0, ref_pres = -1.0;
  bool owns = false;
  FaceCoeffs u{}, v{};
  double *khtr_u = nullptr, *khtr_v = nullptr;
  Workspace ws;
  Timer timer;
  bool is_init = false;
};

extern "C" {

const char* redi_last_error(void) { return g_err.c_str(); }

int redi_version(void) { return 100; }

long long redi_workspace_bytes(int nx, int ny, int nz) {
  if (nx < 1 || ny < 1 || nz < 1) return 0;
  return (long long)(workspace_elems(nx, ny, nz) * sizeof(double));
}

int redi_init(redi_t** out, const redi_hgrid_t* grid, int nz_ml, const redi_coeffs_t* bufs) {
  if (!out || !grid) return fail(REDI_ERR_ARG, "redi_init: null argument");
  *out = nullptr;
  const int nx = grid->nx_total, ny = grid->ny_total;
  if (nx < 1 || ny < 1) return fail(REDI_ERR_ARG, "redi_init: bad grid %d x %d", nx, ny);
  int nz = nz_ml < 1 ? 1 : nz_ml;
  if (nz > REDI_NZ_MAX)
    return fail(REDI_ERR_NZ, "redi_init: 2*nz_ml+2 = %d exceeds the neutral-surface limit %d",
                2 * nz + 2, 2 * REDI_NZ_MAX + 2);
  redi_t* r = new redi_t();
  HIP_TRY(hipGetDevice(&r->device));
  r->nx = nx; r->ny = ny; r->nz = nz; r->ns = 2 * nz + 2;
  const size_t npu = (size_t)(nx + 1) * ny * r->ns, nhu = (size_t)(nx + 1) * ny * (r->ns - 1);
  const size_t npv = (size_t)nx * (ny + 1) * r->ns, nhv = (size_t)nx * (ny + 1) * (r->ns - 1);
  auto cleanup = [&](int code) {
    if (r->owns) {
      hipFree(r->u.PoL); hipFree(r->u.PoR); hipFree(r->u.KoL); hipFree(r->u.KoR); hipFree(r->u.hEff);
      hipFree(r->v.PoL); hipFree(r->v.PoR); hipFree(r->v.KoL); hipFree(r->v.KoR); hipFree(r->v.hEff);
      hipFree(r->khtr_u); hipFree(r->khtr_v);
    }
    hipFree(r->ws.base);
    delete r;
    return code;
  };
  if (bufs) {
    if (!bufs->uPoL || !bufs->uPoR || !bufs->uKoL || !bufs->uKoR || !bufs->uhEff || !bufs->vPoL ||
        !bufs->vPoR || !bufs->vKoL || !bufs->vKoR || !bufs->vhEff || !bufs->khtr_u || !bufs->khtr_v)
      return cleanup(fail(REDI_ERR_ARG, "redi_init: incomplete coefficient buffers"));
    r->u = FaceCoeffs{bufs->uPoL, bufs->uPoR, bufs->uKoL, bufs->uKoR, bufs->uhEff};
    r->v = FaceCoeffs{bufs->vPoL, bufs->vPoR, bufs->vKoL, bufs->vKoR, bufs->vhEff};
    r->khtr_u = bufs->khtr_u;
    r->khtr_v = bufs->khtr_v;
  } else {
    r->owns = true;
    hipError_t e = hipSuccess;
    auto am = [&](void** p, size_t b) { if (e == hipSuccess) e = hipMalloc(p, b); };
    am((void**)&r->u.PoL, npu * 8); am((void**)&r->u.PoR, npu * 8);
    am((void**)&r->u.KoL, npu * 4); am((void**)&r->u.KoR, npu * 4); am((void**)&r->u.hEff, nhu * 8);
    am((void**)&r->v.PoL, npv * 8); am((void**)&r->v.PoR, npv * 8);
    am((void**)&r->v.KoL, npv * 4); am((void**)&r->v.KoR, npv * 4); am((void**)&r->v.hEff, nhv * 8);
    am((void**)&r->khtr_u, (size_t)(nx + 1) * ny * 8);
    am((void**)&r->khtr_v, (size_t)nx * (ny + 1) * 8);
    if (e != hipSuccess)
      return cleanup(fail(REDI_ERR_HIP, "redi_init: hipMalloc: %s", hipGetErrorString(e)));
  }
  {
    const size_t we = workspace_elems(nx, ny, nz);
    hipError_t e = hipMalloc(&r->ws.base, we * sizeof(double));
    if (e != hipSuccess)
      return cleanup(fail(REDI_ERR_HIP, "redi_init: workspace hipMalloc(%zu B): %s",
                          we * sizeof(double), hipGetErrorString(e)));
    r->ws.elems = we;
  }
  // allocate(..., source=0 / 1) (:493-505)
  k_fill_coeffs<<<blocks_for(npu, 256), 256>>>(npu, nhu, r->u);
  k_fill_coeffs<<<blocks_for(npv, 256), 256>>>(npv, nhv, r->v);
  if (hipMemset(r->khtr_u, 0, (size_t)(nx + 1) * ny * 8) != hipSuccess ||
      hipMemset(r->khtr_v, 0, (size_t)nx * (ny + 1) * 8) != hipSuccess ||
      hipDeviceSynchronize() != hipSuccess)
    return cleanup(fail(REDI_ERR_HIP, "redi_init: initial fill failed"));
  r->is_init = true;
  *out = r;
  return REDI_OK;
}

int redi_destroy(redi_t* r) {
  if (!r) return REDI_OK;
  hipDeviceSynchronize();
  if (r->owns) {
    hipFree(r->u.PoL); hipFree(r->u.PoR); hipFree(r->u.KoL); hipFree(r->u.KoR); hipFree(r->u.hEff);
    hipFree(r->v.PoL); hipFree(r->v.PoR); hipFree(r->v.KoL); hipFree(r->v.KoR); hipFree(r->v.hEff);
    hipFree(r->khtr_u); hipFree(r->khtr_v);
  }
  hipFree(r->ws.base);
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

int redi_get_coeffs(const redi_t* r, redi_coeffs_t* o) {
  if (!r || !o) return fail(REDI_ERR_ARG, "redi_get_coeffs: null argument");
  *o = redi_coeffs_t{r->u.PoL, r->u.PoR, r->u.KoL, r->u.KoR, r->u.hEff,
                     r->v.PoL, r->v.PoR, r->v.KoL, r->v.KoR, r->v.hEff,
                     r->khtr_u, r->khtr_v};
  return REDI_OK;
}

int redi_calc_coeffs(const redi_hgrid_t* grid, const redi_metrics_t* m, const redi_eos_t* eos,
                     redi_t* r, const redi_state_t* ms, void* stream) {
  if (!grid || !m || !eos || !ms) return fail(REDI_ERR_ARG, "redi_calc_coeffs: null argument");
  // silent guards (:585-593)
  if (!r || !r->is_init || !r->enable) return REDI_SKIPPED;
  if (ms->idx_temperature <= 0 || ms->idx_salinity <= 0) return REDI_SKIPPED;
  if (!ms->h_layer) return REDI_SKIPPED;
  if (r->nz != ms->nz_ml) return REDI_SKIPPED;
  if (grid->nx_total != r->nx || grid->ny_total != r->ny)
    return fail(REDI_ERR_ARG, "redi_calc_coeffs: grid %dx%d does not match the state %dx%d",
                grid->nx_total, grid->ny_total, r->nx, r->ny);
  if (ms->idx_temperature > ms->ntracers || ms->idx_salinity > ms->ntracers || !ms->tracers)
    return fail(REDI_ERR_ARG, "redi_calc_coeffs: tracer index out of range");
  if (!valid_eos(eos->variant)) return fail(REDI_ERR_ARG, "redi_calc_coeffs: unknown EOS variant %d", eos->variant);
  if (!m->wet_u || !m->wet_v) return fail(REDI_ERR_ARG, "redi_calc_coeffs: null wet mask");
  return launch_phase_a(r->nx, r->ny, r->nz, to_eos(*eos), ms->h_layer,
                        ms->tracers[ms->idx_temperature - 1], ms->tracers[ms->idx_salinity - 1],
                        m->wet_u, m->wet_v, r->u, r->v, r->ws, (hipStream_t)stream, r->timer);
}

int redi_apply_flux(const redi_hgrid_t* grid, const redi_metrics_t* m, redi_t* r,
                    const redi_state_t* ms, double dt, const double* kue, const double* kve,
                    const redi_bc_t* bc, void* stream) {
  if (!grid || !m || !ms) return fail(REDI_ERR_ARG, "redi_apply_flux: null argument");
  // silent guards (:951-959, :978)
  if (!r || !r->is_init || !r->enable) return REDI_SKIPPED;
  if (!ms->h_layer || !ms->tracers || ms->ntracers < 1) return REDI_SKIPPED;
  if (r->nz != ms->nz_ml) return REDI_SKIPPED;
  const bool use_ext = kue && kve;
  if (!use_ext && r->khtr <= 0.0) return REDI_SKIPPED;
  if (grid->nx_total != r->nx || grid->ny_total != r->ny)
    return fail(REDI_ERR_ARG, "redi_apply_flux: grid %dx%d does not match the state %dx%d",
                grid->nx_total, grid->ny_total, r->nx, r->ny);
  if (!m->dy_cu || !m->dx_cv || !m->idxCu || !m->idyCv || !m->areaT)
    return fail(REDI_ERR_ARG, "redi_apply_flux: null metric");
  for (int t = 0; t < ms->ntracers; ++t)
    if (!ms->tracers[t]) return fail(REDI_ERR_ARG, "redi_apply_flux: tracer %d is null", t + 1);
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
  p.khtr_u = r->khtr_u; p.khtr_v = r->khtr_v;
  p.dy_cu = m->dy_cu; p.dx_cv = m->dx_cv; p.idxCu = m->idxCu; p.idyCv = m->idyCv;
  p.areaT = m->areaT; p.h = ms->h_layer;
  p.u = r->u; p.v = r->v;
  return launch_phase_b(p, ms->ntracers, ms->tracers, r->ws, (hipStream_t)stream, r->timer);
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

int redi_launch_nt(int nx, int ny, int nz, double dt, int nghost, int nxp, int nyp, int ww, int we,
                   int wsz, int wn, const redi_eos_t* eos, const double* h_layer, int nt,
                   double* const* tracers, int idx_T, int idx_S, const double* wet_u,
                   const double* wet_v, const redi_coeffs_t* cf, const double* kue,
                   const double* kve, const double* dy_cu, const double* dx_cv, const double* idxCu,
                   const double* idyCv, const double* areaT, void* stream) {
  if (!eos || !cf || !tracers || nt < 1 || nx < 1 || ny < 1 || nz < 1)
    return fail(REDI_ERR_ARG, "redi_launch_nt: bad argument");
  if (nz > REDI_NZ_MAX) return fail(REDI_ERR_NZ, "redi_launch_nt: nz %d > %d", nz, REDI_NZ_MAX);
  if (idx_T < 1 || idx_T > nt || idx_S < 1 || idx_S > nt)
    return fail(REDI_ERR_ARG, "redi_launch_nt: tracer index out of range");
  if (!valid_eos(eos->variant)) return fail(REDI_ERR_ARG, "redi_launch_nt: unknown EOS variant");
  Workspace ws;
  int rc = flat_workspace(workspace_elems(nx, ny, nz), &ws);
  if (rc) return rc;
  const FaceCoeffs u{cf->uPoL, cf->uPoR, cf->uKoL, cf->uKoR, cf->uhEff};
  const FaceCoeffs v{cf->vPoL, cf->vPoR, cf->vKoL, cf->vKoR, cf->vhEff};
  rc = launch_phase_a(nx, ny, nz, to_eos(*eos), h_layer, tracers[idx_T - 1], tracers[idx_S - 1],
                      wet_u, wet_v, u, v, ws, (hipStream_t)stream, g_no_timer);
  if (rc) return rc;
  PhaseB p;
  p.nx = nx; p.ny = ny; p.nz = nz; p.nghost = nghost; p.nxp = nxp; p.nyp = nyp; p.dt = dt;
  p.ww = ww; p.we = we; p.ws = wsz; p.wn = wn;
  p.khtr_u_ext = kue; p.khtr_v_ext = kve; p.khtr = 0.0;
  p.khtr_u = cf->khtr_u; p.khtr_v = cf->khtr_v;
  p.dy_cu = dy_cu; p.dx_cv = dx_cv; p.idxCu = idxCu; p.idyCv = idyCv; p.areaT = areaT;
  p.h = h_layer; p.u = u; p.v = v;
  return launch_phase_b(p, nt, tracers, ws, (hipStream_t)stream, g_no_timer);
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
                                             "k_tracer_recon", "k_apply_flux", "", ""};
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
  hipDeviceSynchronize();
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
