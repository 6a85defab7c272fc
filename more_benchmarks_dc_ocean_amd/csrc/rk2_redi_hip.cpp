// rk2_redi_hip.cpp -- the RK2 benchmark's Redi seam (include/rk2_redi_hip.h)
// over libredi_hip.so.  Host code only; the state lives on the device from
// rk2_redi_init until the process ends.
//
// The seam builds its own analytic state, as the reference's seam does
// (ideal_benchmark/wrap/rk2_redi.F90:77-137, 204-235), so that its probe is
// comparable with the reference's: uniform layers over 4000 m; layer-centre
// temperature 2 + 16 exp(-z/800) + 3 fx + 3 fy and salinity
// 34.5 + 0.7 (1 - exp(-z/1500)) - 0.2 fx, with fx, fy the column's position
// across the domain in [0, 1]; 8 km x 11.1 km metrics; all faces wet;
// khtr = 100 m^2/s; walls on every edge; dt = 900 s.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include <algorithm>
#include <vector>

#include "../../include/redi_hip.h"
#include "../../include/rk2_redi_hip.h"

namespace {

constexpr int kGhost = 3;
constexpr double kDt = 900.0;
constexpr double kDx = 8000.0, kDy = 11100.0, kKappa = 100.0, kDepth = 4000.0;

struct Seam {
  bool ready = false;
  int nxp = 0, nyp = 0, nx = 0, ny = 0, nz = 0;
  redi_t* r = nullptr;
  redi_hgrid_t grid{};
  redi_eos_t eos{};
  redi_bc_t bc{1, 1, 1, 1};
  redi_metrics_t m{};
  redi_state_t st{};
  double* tr[2] = {nullptr, nullptr};
  double *h = nullptr, *ku = nullptr, *kv = nullptr;
  // the reference-layout arrays the flat launchers fill
  double *uPoL, *uPoR, *uhEff, *vPoL, *vPoR, *vhEff, *khtr_u, *khtr_v, *tr_snap;
  int *uKoL, *uKoR, *vKoL, *vKoR;
};
Seam g;

void die(const char* what, int code) {
  fprintf(stderr, "rk2_redi (HIP): %s failed (%d): %s\n", what, code, redi_last_error());
  exit(1);
}

void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) {
    fprintf(stderr, "rk2_redi (HIP): %s: %s\n", what, hipGetErrorString(e));
    exit(1);
  }
}

template <class T>
T* dev_upload(const std::vector<T>& v) {
  T* p = nullptr;
  hip_check(hipMalloc((void**)&p, v.size() * sizeof(T)), "hipMalloc");
  hip_check(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice), "hipMemcpy");
  return p;
}

template <class T>
T* dev_alloc(size_t n) {
  T* p = nullptr;
  hip_check(hipMalloc((void**)&p, n * sizeof(T)), "hipMalloc");
  return p;
}

void stage(bool hoist) {
  if (!g.ready) {
    fprintf(stderr, "rk2_redi (HIP): stage before init\n");
    exit(1);
  }
  int rc = redi_calc_coeffs(&g.grid, &g.m, &g.eos, g.r, &g.st, nullptr);
  if (rc < 0) die("redi_calc_coeffs", rc);
  rc = hoist ? redi_apply_flux_hoist(&g.grid, &g.m, g.r, &g.st, kDt, g.ku, g.kv, &g.bc, nullptr)
             : redi_apply_flux(&g.grid, &g.m, g.r, &g.st, kDt, g.ku, g.kv, &g.bc, nullptr);
  if (rc < 0) die("redi_apply_flux", rc);
  hip_check(hipDeviceSynchronize(), "stage");
}

template <class Launch>
void stage_flat(Launch launch) {
  if (!g.ready) {
    fprintf(stderr, "rk2_redi (HIP): stage before init\n");
    exit(1);
  }
  const int ns = 2 * g.nz + 2, ng = kGhost, one = 1;
  const double dt = kDt;
  launch(&g.nx, &g.ny, &g.nz, &ns, &dt, &ng, &g.nxp, &g.nyp, &one, &one, &one, &one, &g.eos.rho0,
         &g.eos.T_ref, &g.eos.S_ref, &g.eos.alpha_T, &g.eos.beta_S, g.h, g.tr[0], g.tr[1],
         g.m.wet_u, g.m.wet_v, g.uPoL, g.uPoR, g.uKoL, g.uKoR, g.uhEff, g.vPoL, g.vPoR, g.vKoL,
         g.vKoR, g.vhEff, g.ku, g.kv, g.khtr_u, g.khtr_v, g.m.dy_cu, g.m.dx_cv, g.m.idxCu,
         g.m.idyCv, g.m.areaT, g.tr_snap);
  hip_check(hipGetLastError(), "flat launch");
}

}  // namespace

extern "C" {

void rk2_redi_init(int nxp, int nyp, int nz) {
  if (g.ready) return;
  if (nxp < 2 || nyp < 2 || nz < 1) {
    fprintf(stderr, "rk2_redi (HIP): bad size %d x %d x %d\n", nxp, nyp, nz);
    exit(1);
  }
  const int nx = nxp + 2 * kGhost, ny = nyp + 2 * kGhost;
  g.nxp = nxp; g.nyp = nyp; g.nx = nx; g.ny = ny; g.nz = nz;
  g.grid = redi_hgrid_t{nx, ny, nxp, nyp, kGhost, 0, 0.1, 0.1};
  g.eos = redi_eos_t{REDI_EOS_LINEAR, 1, 1035.0, 10.0, 35.0, 1.7e-4, 7.6e-4, 0.0};
  const size_t n2 = (size_t)nx * ny, n3 = n2 * nz;
  const size_t nu = (size_t)(nx + 1) * ny, nv = (size_t)nx * (ny + 1);
  // the analytic state, native layout (k = 0 at the bed)
  std::vector<double> h(n3), t(n3), s(n3);
  const double hh = kDepth / (double)nz;
  for (int k = 0; k < nz; ++k) {
    const double z = ((double)(nz - 1 - k) + 0.5) * hh;  // layer-centre depth
    for (int j = 0; j < ny; ++j) {
      const double fy = (double)j / (double)(ny - 1);
      for (int i = 0; i < nx; ++i) {
        const double fx = (double)i / (double)(nx - 1);
        const double tt = 2.0 + 16.0 * exp(-z / 800.0) + 3.0 * fx + 3.0 * fy;
        const double ss = 34.5 + 0.7 * (1.0 - exp(-z / 1500.0)) - 0.2 * fx;
        const size_t o = (size_t)k * n2 + (size_t)j * nx + i;
        h[o] = hh;
        t[o] = tt * hh;
        s[o] = ss * hh;
      }
    }
  }
  g.h = dev_upload(h);
  g.tr[0] = dev_upload(t);
  g.tr[1] = dev_upload(s);
  g.m.dy_cu = dev_upload(std::vector<double>(nu, kDy));
  g.m.dx_cv = dev_upload(std::vector<double>(nv, kDx));
  g.m.idxCu = dev_upload(std::vector<double>(nu, 1.0 / kDx));
  g.m.idyCv = dev_upload(std::vector<double>(nv, 1.0 / kDy));
  g.m.areaT = dev_upload(std::vector<double>(n2, kDx * kDy));
  g.m.wet_u = dev_upload(std::vector<double>(nu, 1.0));
  g.m.wet_v = dev_upload(std::vector<double>(nv, 1.0));
  g.ku = dev_upload(std::vector<double>(nu, kKappa));
  g.kv = dev_upload(std::vector<double>(nv, kKappa));
  const size_t ns = 2 * (size_t)nz + 2;
  g.uPoL = dev_alloc<double>(nu * ns); g.uPoR = dev_alloc<double>(nu * ns);
  g.uKoL = dev_alloc<int>(nu * ns); g.uKoR = dev_alloc<int>(nu * ns);
  g.uhEff = dev_alloc<double>(nu * (ns - 1));
  g.vPoL = dev_alloc<double>(nv * ns); g.vPoR = dev_alloc<double>(nv * ns);
  g.vKoL = dev_alloc<int>(nv * ns); g.vKoR = dev_alloc<int>(nv * ns);
  g.vhEff = dev_alloc<double>(nv * (ns - 1));
  g.khtr_u = dev_alloc<double>(nu); g.khtr_v = dev_alloc<double>(nv);
  g.tr_snap = dev_alloc<double>(n3);
  g.st = redi_state_t{nz, 1, 2, 2, g.h, g.tr};
  int rc = redi_init(&g.r, &g.grid, nz);
  if (rc < 0) die("redi_init", rc);
  rc = redi_configure(g.r, 1, kKappa, 1, -1.0);
  if (rc < 0) die("redi_configure", rc);
  g.ready = true;
  rc = redi_calc_coeffs(&g.grid, &g.m, &g.eos, g.r, &g.st, nullptr);
  if (rc < 0) die("redi_calc_coeffs", rc);
  hip_check(hipDeviceSynchronize(), "init");
}

void rk2_redi_stage(void) { stage(true); }
void rk2_redi_stage_unopt(void) { stage(false); }
void rk2_redi_stage_cuda(void) { stage_flat(redi_opt_launch); }
void rk2_redi_stage_cuda_unopt(void) { stage_flat(redi_cuda_launch_); }

void rk2_redi_probe(double* vmin, double* vmax) {
  if (!g.ready) {
    *vmin = *vmax = 0.0;
    return;
  }
  std::vector<double> t((size_t)g.nx * g.ny * g.nz);
  hip_check(hipMemcpy(t.data(), g.tr[0], t.size() * sizeof(double), hipMemcpyDeviceToHost), "probe");
  const auto mm = std::minmax_element(t.begin(), t.end());
  *vmin = *mm.first;
  *vmax = *mm.second;
}

int rk2_redi_export(double* h, double* t_htr, double* s_htr) {
  if (!g.ready) return -1;
  const size_t b = (size_t)g.nx * g.ny * g.nz * sizeof(double);
  hip_check(hipMemcpy(h, g.h, b, hipMemcpyDeviceToHost), "export h");
  hip_check(hipMemcpy(t_htr, g.tr[0], b, hipMemcpyDeviceToHost), "export t");
  hip_check(hipMemcpy(s_htr, g.tr[1], b, hipMemcpyDeviceToHost), "export s");
  return 0;
}

}  // extern "C"
