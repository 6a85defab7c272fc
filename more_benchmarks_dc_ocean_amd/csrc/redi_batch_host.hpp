// redi_batch_host.hpp -- host side of the batched form (include/redi_hip.h,
// redi_batch_*).  Part of redi_lib.hip's translation unit: it builds the
// same kernel arguments as the single-domain launches (launch_phase_a,
// phase_b_args) and hands them to the persistent kernels of
// redi_batch_kernels.hpp through device-resident work tables.
//
//   arena    one hipMalloc of the batch's HBM budget at create; each domain's
//            Phase-A face state and step scratch is a first-fit range of it,
//            so admitting or retiring a domain allocates nothing on the device
//   tables   per kernel kind: item count and virtual block count, each
//            item's first virtual block (a multiple of 8) and block count, and
//            the items (the kernel's argument struct of one domain launch);
//            built in pinned host memory and copied to the device in stream
//            order before the first step after a change
//   graph    the step's launches -- column state, sweep, face diffusivities,
//            reconstruction, x and y fluxes per layer-capacity class, tracer
//            update --
//            captured once into a hipGraph: the persistent grids do not depend
//            on the batch, so table changes never invalidate it
#pragma once

#include <map>

#include "redi_batch_kernels.hpp"

namespace {

inline size_t round256(size_t n) { return (n + 255) & ~(size_t)255; }

// First-fit allocator over [0, cap) with coalescing of freed ranges.
class Arena {
 public:
  void reset(size_t cap) {
    cap_ = cap;
    used_ = 0;
    free_.clear();
    if (cap) free_[0] = cap;
  }
  bool take(size_t bytes, size_t* off) {
    bytes = round256(bytes);
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second < bytes) continue;
      *off = it->first;
      const size_t rest = it->second - bytes, o2 = it->first + bytes;
      free_.erase(it);
      if (rest) free_[o2] = rest;
      used_ += bytes;
      return true;
    }
    return false;
  }
  void give(size_t off, size_t bytes) {
    bytes = round256(bytes);
    used_ -= bytes;
    auto it = free_.emplace(off, bytes).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free_.erase(nx);
    }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_.erase(it);
      }
    }
  }
  size_t used() const { return used_; }
  size_t largest() const {
    size_t m = 0;
    for (const auto& kv : free_) m = std::max(m, kv.second);
    return m;
  }

 private:
  size_t cap_ = 0, used_ = 0;
  std::map<size_t, size_t> free_;  // offset -> bytes
};

// One domain's ranges inside its arena block (offsets from the block start).
struct DomainLayout {
  size_t u, v, mono_u, mono_v, ku, kv, ws, total;
};

DomainLayout domain_layout(int nx, int ny, int nz, int form, int group) {
  const size_t ns = 2 * (size_t)nz + 2, nu = (size_t)(nx + 1) * ny, nv = (size_t)nx * (ny + 1);
  const size_t fb = sizes_of(form).face;
  DomainLayout L;
  size_t off = 0;
  auto take = [&](size_t bytes) {
    const size_t o = off;
    off += round256(bytes);
    return o;
  };
  L.u = take(padded_faces(nu) * ns * fb);
  L.v = take(padded_faces(nv) * ns * fb);
  L.mono_u = take(nu);
  L.mono_v = take(nv);
  L.ku = take(nu * sizeof(double));
  L.kv = take(nv * sizeof(double));
  L.ws = take(workspace_bytes(nx, ny, nz, group, form));
  L.total = off;
  return L;
}

enum BatchNode {
  kNodeColumn, kNodeSweep, kNodeKhtr, kNodeRecon,
  kNodeFluxX0, kNodeFluxY0 = kNodeFluxX0 + 4, kNodeUpdate = kNodeFluxY0 + 4, kNodes
};

struct NodeTable {
  size_t count = 0, start = 0, nblk = 0, items = 0;  // byte offsets in the table buffer
  size_t item_size = 0;
  int cap = 0;  // items
};

struct BatchDomain {
  bool live = false;
  int nx = 0, ny = 0, nz = 0, nt = 0, idx_t = 0, idx_s = 0;
  const void* h = nullptr;
  std::vector<void*> tr;
  const double *wet_u = nullptr, *wet_v = nullptr;
  EosParams eos{};
  double slope_max = 0.0;
  PhaseB p{};
  Tiles tiles;
  FaceState fs;
  void* ws = nullptr;
  size_t off = 0, bytes = 0;
};

}  // namespace

struct redi_batch_s {
  int device = 0, form = REDI_FORM_FP64, group = kTracerGroup, max_domains = 0, persist = 1;
  int ncu = 0;
  char* arena_base = nullptr;
  Arena arena;
  std::vector<BatchDomain> dom;
  int ndom = 0;
  NodeTable tab[kNodes];
  size_t tab_bytes = 0;
  char* dtab = nullptr;  // device tables
  char* htab = nullptr;  // pinned host staging
  hipEvent_t uploaded = nullptr;
  bool upload_pending = false;  // an upload from htab may still be in flight
  bool dirty = true;            // tables differ from the domains
  hipStream_t cap_stream = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  bool graph_stale = true;
  unsigned grid[kNodes] = {};
  uint32_t total[kNodes] = {};  // virtual blocks of the current tables
};

namespace {

template <class F, int NT>
struct BatchOps {
  using R = typename F::R;
  using Recon = ReconItem<F, NT>;

  static size_t item_size(int node) {
    switch (node) {
      case kNodeColumn: return sizeof(ColumnArgs<F>);
      case kNodeSweep: return sizeof(SweepItem<F>);
      case kNodeKhtr: return sizeof(KhtrArgs);
      case kNodeRecon: return sizeof(Recon);
      case kNodeUpdate: return sizeof(UpdateArgs<F>);
      default: return sizeof(FluxArgs<F>);
    }
  }

  template <class Item>
  static BatchView<Item> view(const redi_batch_s* b, int node) {
    const NodeTable& t = b->tab[node];
    return BatchView<Item>{(const uint32_t*)(b->dtab + t.count), (const uint32_t*)(b->dtab + t.start),
                           (const uint32_t*)(b->dtab + t.nblk), (const Item*)(b->dtab + t.items)};
  }

  // One node's table in the pinned staging buffer.
  struct Writer {
    redi_batch_s* b;
    int node;
    uint32_t n = 0, next = 0;
    template <class Item>
    void add(const Item& it, uint32_t blocks) {
      if (!blocks) return;
      const NodeTable& t = b->tab[node];
      uint32_t* start = (uint32_t*)(b->htab + t.start);
      uint32_t* nblk = (uint32_t*)(b->htab + t.nblk);
      start[n] = next;
      nblk[n] = blocks;
      memcpy(b->htab + t.items + (size_t)n * t.item_size, &it, sizeof(Item));
      next += (blocks + 7) & ~7u;  // the next item starts on a multiple of 8
      ++n;
    }
    void finish() {
      const NodeTable& t = b->tab[node];
      uint32_t* cnt = (uint32_t*)(b->htab + t.count);
      uint32_t* start = (uint32_t*)(b->htab + t.start);
      start[n] = next;
      cnt[0] = n;
      cnt[1] = next;
      b->total[node] = next;
    }
  };

  static void build(redi_batch_s* b) {
    Writer w[kNodes];
    for (int k = 0; k < kNodes; ++k) w[k] = Writer{b, k};
    for (const BatchDomain& d : b->dom) {
      if (!d.live) continue;
      const int nx = d.nx, ny = d.ny, nz = d.nz, N2 = nx * ny;
      const R* h = (const R*)d.h;
      R* const* tr = (R* const*)d.tr.data();
      // Phase A (launch_phase_a)
      typename F::Col* rec = (typename F::Col*)d.ws;
      w[kNodeColumn].add(ColumnArgs<F>{N2, nz, nx, h, tr[d.idx_t - 1], tr[d.idx_s - 1], d.eos, rec},
                         blocks_for(N2, 128));
      const double pa_to_h = 1.0 / (kGravity * d.eos.rho0);
      if (nx > 1) {
        SweepItem<F> it{SweepArgs<F>{nx, nz, nx + 1, face_fa(2 * nz + 2), face_fb(2 * nz + 2), 1, 0, 1, 0,
                                     d.wet_u, rec,
                                     (FaceRecF<F>*)d.fs.u, d.fs.mono_u, pa_to_h, d.p.idxCu, d.slope_max,
                                     d.fs.filled ? 1 : 0},
                        TileMap{}};
        unsigned nb;
        it.tm = make_tiles(nx - 1, ny, d.tiles, &nb);
        w[kNodeSweep].add(it, nb);
      }
      if (ny > 1) {
        SweepItem<F> it{SweepArgs<F>{nx, nz, nx, face_fa(2 * nz + 2), face_fb(2 * nz + 2), 0, 1, 0, 1,
                                     d.wet_v, rec,
                                     (FaceRecF<F>*)d.fs.v, d.fs.mono_v, pa_to_h, d.p.idyCv, d.slope_max,
                                     d.fs.filled ? 1 : 0},
                        TileMap{}};
        unsigned nb;
        it.tm = make_tiles(nx, ny - 1, d.tiles, &nb);
        w[kNodeSweep].add(it, nb);
      }
      // Phase B (launch_phase_b with one tracer group of NT slots)
      const PhaseB& p = d.p;
      const int nu = (nx + 1) * ny, nv = nx * (ny + 1);
      w[kNodeKhtr].add(KhtrArgs{nu, nv, p.khtr_u_ext, p.khtr_v_ext, p.khtr, d.fs.khtr_u, d.fs.khtr_v},
                       blocks_for((size_t)nu + nv, 256));
      const PhaseBScratch sc = carve_phase_b(d.ws, nx, ny, nz, NT, form_sizes<F>());
      FluxArgs<F> fx, fy;
      UpdateArgs<F> ua;
      unsigned nbx = 0, nby = 0;
      phase_b_args<F>(p, d.fs, sc, d.tiles.seg, fx, fy, ua, &nbx, &nby);
      TracerSetT<R> ts{};
      for (int t = 0; t < kMaxNT; ++t) {
        // slots past the domain's tracers repeat its last tracer; their
        // results land in the discard field
        ts.q[t] = tr[std::min(t, d.nt - 1)];
        ua.htr[t] = t < d.nt ? tr[t] : (R*)sc.discard;
        ua.tend[t] = nullptr;
      }
      ua.nt = NT;
      const ReconLayout ry = recon_layout(0, N2, nz + F::Recon::kExtra, NT);
      fx.ry = fy.ry = ry;
      w[kNodeRecon].add(Recon{ReconArgs<F, NT>{N2, nz, h, ts, (typename F::Recon*)sc.recon, ry}, d.tiles.recon_kt, 0},
                        recon_tiled_blocks(N2, nz, d.tiles.recon_kt));
      const int fc = flux_class(nz);
      w[kNodeFluxX0 + fc].add(fx, nbx);
      w[kNodeFluxY0 + fc].add(fy, nby);
      w[kNodeUpdate].add(ua, update_blocks(N2, nz));
    }
    for (int k = 0; k < kNodes; ++k) w[k].finish();
  }

  template <class K>
  static unsigned persistent_grid(const redi_batch_s* b, K kernel, int bs, int node) {
    if (b->persist <= 0) return std::max(1u, b->total[node]);
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, kernel, bs, 0) != hipSuccess || n < 1) n = 1;
    const unsigned g = (unsigned)n * (unsigned)b->ncu * (unsigned)b->persist;
    return std::max(8u, g / 8 * 8);
  }

  static void grids(redi_batch_s* b) {
    b->grid[kNodeColumn] = persistent_grid(b, kb_column_state<F>, 128, kNodeColumn);
    b->grid[kNodeSweep] = persistent_grid(b, kb_neutral_sweep<F>, kBatchSweepLanes, kNodeSweep);
    b->grid[kNodeKhtr] = persistent_grid(b, kb_face_khtr, 256, kNodeKhtr);
    b->grid[kNodeRecon] = persistent_grid(b, kb_tracer_recon<F, NT>, kReconLanes, kNodeRecon);
    b->grid[kNodeFluxX0 + 0] = persistent_grid(b, kb_flux<F, NT, kFluxNzc[0], false>, 64, kNodeFluxX0 + 0);
    b->grid[kNodeFluxX0 + 1] = persistent_grid(b, kb_flux<F, NT, kFluxNzc[1], false>, 64, kNodeFluxX0 + 1);
    b->grid[kNodeFluxX0 + 2] = persistent_grid(b, kb_flux<F, NT, kFluxNzc[2], false>, 64, kNodeFluxX0 + 2);
    b->grid[kNodeFluxX0 + 3] = persistent_grid(b, kb_flux<F, NT, kFluxNzc[3], false>, 64, kNodeFluxX0 + 3);
    b->grid[kNodeFluxY0 + 0] = persistent_grid(b, kb_flux<F, NT, kFluxNzc[0], true>, 64, kNodeFluxY0 + 0);
    b->grid[kNodeFluxY0 + 1] = persistent_grid(b, kb_flux<F, NT, kFluxNzc[1], true>, 64, kNodeFluxY0 + 1);
    b->grid[kNodeFluxY0 + 2] = persistent_grid(b, kb_flux<F, NT, kFluxNzc[2], true>, 64, kNodeFluxY0 + 2);
    b->grid[kNodeFluxY0 + 3] = persistent_grid(b, kb_flux<F, NT, kFluxNzc[3], true>, 64, kNodeFluxY0 + 3);
    b->grid[kNodeUpdate] = persistent_grid(b, kb_tracer_update<F>, kUpdLanes, kNodeUpdate);
  }

  // The step's launches on `st` (captured into the graph).  A kernel kind
  // with no work in the batch is left out.
  static void launch_all(const redi_batch_s* b, hipStream_t st) {
    const unsigned* g = b->grid;
    auto live = [&](int node) { return b->persist > 0 || b->total[node] > 0; };
    if (live(kNodeColumn))
      kb_column_state<F><<<g[kNodeColumn], 128, 0, st>>>(view<ColumnArgs<F>>(b, kNodeColumn));
    if (live(kNodeSweep))
      kb_neutral_sweep<F><<<g[kNodeSweep], kBatchSweepLanes, 0, st>>>(view<SweepItem<F>>(b, kNodeSweep));
    if (live(kNodeKhtr)) kb_face_khtr<<<g[kNodeKhtr], 256, 0, st>>>(view<KhtrArgs>(b, kNodeKhtr));
    if (live(kNodeRecon))
      kb_tracer_recon<F, NT><<<g[kNodeRecon], kReconLanes, 0, st>>>(view<Recon>(b, kNodeRecon));
    using FA = FluxArgs<F>;
    // every x class before any y class: the y kernels read the dTr partials
    // the x kernels wrote (stream order)
    if (live(kNodeFluxX0 + 0))
      kb_flux<F, NT, kFluxNzc[0], false><<<g[kNodeFluxX0 + 0], 64, 0, st>>>(view<FA>(b, kNodeFluxX0 + 0));
    if (live(kNodeFluxX0 + 1))
      kb_flux<F, NT, kFluxNzc[1], false><<<g[kNodeFluxX0 + 1], 64, 0, st>>>(view<FA>(b, kNodeFluxX0 + 1));
    if (live(kNodeFluxX0 + 2))
      kb_flux<F, NT, kFluxNzc[2], false><<<g[kNodeFluxX0 + 2], 64, 0, st>>>(view<FA>(b, kNodeFluxX0 + 2));
    if (live(kNodeFluxX0 + 3))
      kb_flux<F, NT, kFluxNzc[3], false><<<g[kNodeFluxX0 + 3], 64, 0, st>>>(view<FA>(b, kNodeFluxX0 + 3));
    if (live(kNodeFluxY0 + 0))
      kb_flux<F, NT, kFluxNzc[0], true><<<g[kNodeFluxY0 + 0], 64, 0, st>>>(view<FA>(b, kNodeFluxY0 + 0));
    if (live(kNodeFluxY0 + 1))
      kb_flux<F, NT, kFluxNzc[1], true><<<g[kNodeFluxY0 + 1], 64, 0, st>>>(view<FA>(b, kNodeFluxY0 + 1));
    if (live(kNodeFluxY0 + 2))
      kb_flux<F, NT, kFluxNzc[2], true><<<g[kNodeFluxY0 + 2], 64, 0, st>>>(view<FA>(b, kNodeFluxY0 + 2));
    if (live(kNodeFluxY0 + 3))
      kb_flux<F, NT, kFluxNzc[3], true><<<g[kNodeFluxY0 + 3], 64, 0, st>>>(view<FA>(b, kNodeFluxY0 + 3));
    if (live(kNodeUpdate))
      kb_tracer_update<F><<<g[kNodeUpdate], kUpdLanes, 0, st>>>(view<UpdateArgs<F>>(b, kNodeUpdate));
  }

  static int capture(redi_batch_s* b) {
    if (b->exec) { (void)hipGraphExecDestroy(b->exec); b->exec = nullptr; }
    if (b->graph) { (void)hipGraphDestroy(b->graph); b->graph = nullptr; }
    grids(b);
    HIP_TRY(hipStreamBeginCapture(b->cap_stream, hipStreamCaptureModeThreadLocal));
    launch_all(b, b->cap_stream);
    const hipError_t le = hipGetLastError();
    hipGraph_t g = nullptr;
    const hipError_t ce = hipStreamEndCapture(b->cap_stream, &g);
    if (le != hipSuccess || ce != hipSuccess) {
      if (g) (void)hipGraphDestroy(g);
      return fail(REDI_ERR_HIP, "redi_batch_step: graph capture failed: %s",
                  hipGetErrorString(le != hipSuccess ? le : ce));
    }
    b->graph = g;
    HIP_TRY(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    b->graph_stale = false;
    return REDI_OK;
  }

  static int serial(redi_batch_s* b, hipStream_t st) {
    for (BatchDomain& d : b->dom) {
      if (!d.live) continue;
      R* const* tr = (R* const*)d.tr.data();
      int rc = launch_phase_a<F>(d.nx, d.ny, d.nz, d.eos, (const R*)d.h, tr[d.idx_t - 1],
                                 tr[d.idx_s - 1], d.wet_u, d.wet_v, d.fs, d.ws, d.tiles, st, g_no_timer,
                                 Taper{d.p.idxCu, d.p.idyCv, d.slope_max});
      if (rc) return rc;
      rc = launch_phase_b<F>(d.p, (const R*)d.h, d.nt, tr, d.fs, d.ws, d.tiles, st, g_no_timer);
      if (rc) return rc;
    }
    return REDI_OK;
  }
};

template <class Fn>
int with_batch_ops(int form, int group, Fn&& fn) {
  if (form == REDI_FORM_FP32_STORE) {
    switch (group) {
      case 2: return fn(BatchOps<Fp32, 2>{});
      case 3: return fn(BatchOps<Fp32, 3>{});
      default: return fn(BatchOps<Fp32, 4>{});
    }
  }
  switch (group) {
    case 2: return fn(BatchOps<Fp64, 2>{});
    case 3: return fn(BatchOps<Fp64, 3>{});
    default: return fn(BatchOps<Fp64, 4>{});
  }
}

void batch_free(redi_batch_s* b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  if (b->cap_stream) (void)hipStreamDestroy(b->cap_stream);
  if (b->uploaded) (void)hipEventDestroy(b->uploaded);
  if (b->htab) (void)hipHostFree(b->htab);
  if (b->dtab) (void)hipFree(b->dtab);
  if (b->arena_base) (void)hipFree(b->arena_base);
  delete b;
}

// Rebuild the tables in the staging buffer and copy them to the device on
// `st` (after any earlier copy from the staging buffer has finished).
int batch_upload(redi_batch_s* b, hipStream_t st) {
  if (b->upload_pending) {
    HIP_TRY(hipEventSynchronize(b->uploaded));
    b->upload_pending = false;
  }
  uint32_t old_total[kNodes];
  memcpy(old_total, b->total, sizeof old_total);
  with_batch_ops(b->form, b->group, [&](auto ops) {
    decltype(ops)::build(b);
    return 0;
  });
  HIP_TRY(hipMemcpyAsync(b->dtab, b->htab, b->tab_bytes, hipMemcpyHostToDevice, st));
  HIP_TRY(hipEventRecord(b->uploaded, st));
  b->upload_pending = true;
  b->dirty = false;
  if (b->persist <= 0) {  // grids follow the work: re-capture when it changed
    for (int k = 0; k < kNodes; ++k)
      if (old_total[k] != b->total[k]) b->graph_stale = true;
  }
  return REDI_OK;
}

int check_domain(const redi_batch_s* b, const redi_domain_t* d) {
  const int nx = d->grid.nx_total, ny = d->grid.ny_total, nz = d->nz_ml;
  if (nx < 1 || ny < 1) return fail(REDI_ERR_ARG, "redi_batch_add: bad grid %d x %d", nx, ny);
  if (nz < 1 || nz > REDI_NZ_MAX)
    return fail(REDI_ERR_NZ, "redi_batch_add: nz %d not in 1..%d", nz, REDI_NZ_MAX);
  if ((double)(nz + 1) * nx * ny * kTracerGroup >= 2147483648.0)
    return fail(REDI_ERR_ARG, "redi_batch_add: %d x %d x %d exceeds 2^29 cell interfaces", nx, ny, nz);
  if (d->ntracers < 1 || d->ntracers > b->group)
    return fail(REDI_ERR_ARG, "redi_batch_add: %d tracers; this batch holds 1..%d per domain",
                d->ntracers, b->group);
  if (d->idx_temperature < 1 || d->idx_salinity < 1 || d->idx_temperature > d->ntracers ||
      d->idx_salinity > d->ntracers)
    return fail(REDI_ERR_ARG, "redi_batch_add: tracer index out of range");
  if (!valid_eos(d->eos.variant))
    return fail(REDI_ERR_ARG, "redi_batch_add: unknown EOS variant %d", d->eos.variant);
  const redi_metrics_t& m = d->metrics;
  if (!m.dy_cu || !m.dx_cv || !m.idxCu || !m.idyCv || !m.areaT || !m.wet_u || !m.wet_v)
    return fail(REDI_ERR_ARG, "redi_batch_add: null metric");
  if (!d->h_layer || !d->tracers) return fail(REDI_ERR_ARG, "redi_batch_add: null state");
  for (int t = 0; t < d->ntracers; ++t)
    if (!d->tracers[t]) return fail(REDI_ERR_ARG, "redi_batch_add: tracer %d is null", t + 1);
  const bool use_ext = d->khtr_u_ext && d->khtr_v_ext;
  if (!use_ext && !(d->khtr > 0.0))
    return fail(REDI_ERR_ARG, "redi_batch_add: no face diffusivity (khtr <= 0 and no khtr_*_ext)");
  const int ti = d->tile_i ? d->tile_i : kBatchSweepLanes, tj = d->tile_j ? d->tile_j : 1;
  if (ti < 1 || tj < 1 || ti * tj != kBatchSweepLanes)
    return fail(REDI_ERR_ARG, "redi_batch_add: sweep tile %d x %d must hold %d lanes", ti, tj,
                kBatchSweepLanes);
  if (!(d->slope_max >= 0.0)) return fail(REDI_ERR_ARG, "redi_batch_add: slope limit %g < 0", d->slope_max);
  if (d->kchunk && d->kchunk != 8 && d->kchunk != 16 && d->kchunk != 32)
    return fail(REDI_ERR_ARG, "redi_batch_add: layer chunk %d is not 8, 16 or 32", d->kchunk);
  return REDI_OK;
}

}  // namespace

extern "C" {

long long redi_batch_domain_bytes(int nx, int ny, int nz, int form, int group) {
  if (nx < 1 || ny < 1 || nz < 1 || group < 2 || group > kTracerGroup) return 0;
  if (form != REDI_FORM_FP64 && form != REDI_FORM_FP32_STORE) return 0;
  return (long long)domain_layout(nx, ny, nz, form, group).total;
}

int redi_batch_create(redi_batch_t** out, int form, int group, int max_domains, long long arena_bytes,
                      int persist) {
  if (!out) return fail(REDI_ERR_ARG, "redi_batch_create: null argument");
  *out = nullptr;
  if (form != REDI_FORM_FP64 && form != REDI_FORM_FP32_STORE)
    return fail(REDI_ERR_ARG, "redi_batch_create: unknown form %d", form);
  if (group < 2 || group > kTracerGroup)
    return fail(REDI_ERR_ARG, "redi_batch_create: group %d not in 2..%d", group, kTracerGroup);
  if (max_domains < 1 || max_domains > 65536)
    return fail(REDI_ERR_ARG, "redi_batch_create: max_domains %d not in 1..65536", max_domains);
  if (arena_bytes < 1) return fail(REDI_ERR_ARG, "redi_batch_create: empty arena");
  if (persist < 0 || persist > 64) return fail(REDI_ERR_ARG, "redi_batch_create: persist %d not in 0..64", persist);
  redi_batch_s* b = new redi_batch_s();
  b->form = form;
  b->group = group;
  b->max_domains = max_domains;
  b->persist = persist;
  auto bail = [&](int code) {
    batch_free(b);
    return code;
  };
  if (hipGetDevice(&b->device) != hipSuccess) return bail(fail(REDI_ERR_HIP, "redi_batch_create: no device"));
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, b->device) != hipSuccess)
    return bail(fail(REDI_ERR_HIP, "redi_batch_create: device properties"));
  b->ncu = prop.multiProcessorCount;
  if (hipMalloc((void**)&b->arena_base, (size_t)arena_bytes) != hipSuccess)
    return bail(fail(REDI_ERR_HIP, "redi_batch_create: arena hipMalloc(%lld B) failed", arena_bytes));
  b->arena.reset((size_t)arena_bytes);
  b->dom.resize(max_domains);
  // table layout: per node count[2], start[cap + 1], nblk[cap], items[cap]
  size_t off = 0;
  for (int k = 0; k < kNodes; ++k) {
    NodeTable& t = b->tab[k];
    t.cap = k == kNodeSweep ? 2 * max_domains : max_domains;
    t.item_size = (size_t)with_batch_ops(form, group, [&](auto ops) {
      return (int)decltype(ops)::item_size(k);
    });
    t.count = off; off = round256(off + 2 * sizeof(uint32_t));
    t.start = off; off = round256(off + (t.cap + 1) * sizeof(uint32_t));
    t.nblk = off; off = round256(off + t.cap * sizeof(uint32_t));
    t.items = off; off = round256(off + t.cap * t.item_size);
  }
  b->tab_bytes = off;
  if (hipMalloc((void**)&b->dtab, off) != hipSuccess)
    return bail(fail(REDI_ERR_HIP, "redi_batch_create: table hipMalloc failed"));
  if (hipHostMalloc((void**)&b->htab, off, hipHostMallocDefault) != hipSuccess)
    return bail(fail(REDI_ERR_HIP, "redi_batch_create: pinned staging failed"));
  memset(b->htab, 0, off);
  if (hipEventCreateWithFlags(&b->uploaded, hipEventDisableTiming) != hipSuccess ||
      hipStreamCreateWithFlags(&b->cap_stream, hipStreamNonBlocking) != hipSuccess)
    return bail(fail(REDI_ERR_HIP, "redi_batch_create: event / stream"));
  // empty tables on the device before any launch reads them
  if (hipMemcpy(b->dtab, b->htab, off, hipMemcpyHostToDevice) != hipSuccess)
    return bail(fail(REDI_ERR_HIP, "redi_batch_create: table upload failed"));
  *out = b;
  return REDI_OK;
}

int redi_batch_destroy(redi_batch_t* b) {
  if (!b) return REDI_OK;
  (void)hipDeviceSynchronize();
  batch_free(b);
  return REDI_OK;
}

int redi_batch_add(redi_batch_t* b, const redi_domain_t* d, void* stream) {
  if (!b || !d) return fail(REDI_ERR_ARG, "redi_batch_add: null argument");
  int rc = check_domain(b, d);
  if (rc) return rc;
  int slot = -1;
  for (int s = 0; s < b->max_domains; ++s)
    if (!b->dom[s].live) { slot = s; break; }
  if (slot < 0) return fail(REDI_ERR_BUDGET, "redi_batch_add: all %d domain slots are in use", b->max_domains);
  const int nx = d->grid.nx_total, ny = d->grid.ny_total, nz = d->nz_ml;
  const DomainLayout L = domain_layout(nx, ny, nz, b->form, b->group);
  size_t off = 0;
  if (!b->arena.take(L.total, &off))
    return fail(REDI_ERR_BUDGET, "redi_batch_add: %zu B do not fit the arena (largest free range %zu B)",
                L.total, b->arena.largest());
  BatchDomain& x = b->dom[slot];
  x = BatchDomain{};
  x.nx = nx; x.ny = ny; x.nz = nz; x.nt = d->ntracers;
  x.idx_t = d->idx_temperature; x.idx_s = d->idx_salinity;
  x.h = d->h_layer;
  x.tr.assign(d->tracers, d->tracers + d->ntracers);
  x.wet_u = d->metrics.wet_u; x.wet_v = d->metrics.wet_v;
  x.eos = to_eos(d->eos);
  x.slope_max = d->slope_max;
  PhaseB& p = x.p;
  p.nx = nx; p.ny = ny; p.nz = nz;
  p.nghost = d->grid.nghost; p.nxp = d->grid.nx_phys; p.nyp = d->grid.ny_phys;
  p.dt = d->dt;
  p.ww = d->bc.west == REDI_OBC_WALL; p.we = d->bc.east == REDI_OBC_WALL;
  p.ws = d->bc.south == REDI_OBC_WALL; p.wn = d->bc.north == REDI_OBC_WALL;
  const bool use_ext = d->khtr_u_ext && d->khtr_v_ext;
  p.khtr_u_ext = use_ext ? d->khtr_u_ext : nullptr;
  p.khtr_v_ext = use_ext ? d->khtr_v_ext : nullptr;
  p.khtr = d->khtr;
  p.dy_cu = d->metrics.dy_cu; p.dx_cv = d->metrics.dx_cv;
  p.idxCu = d->metrics.idxCu; p.idyCv = d->metrics.idyCv; p.areaT = d->metrics.areaT;
  x.tiles.tx = d->tile_i ? d->tile_i : kBatchSweepLanes;
  x.tiles.ty = d->tile_j ? d->tile_j : 1;
  x.tiles.recon_kt = d->kchunk ? d->kchunk : kReconChunkDefault;
  x.tiles.group = b->group;
  x.tiles.recon_layer_major = 0;
  char* base = b->arena_base + off;
  x.off = off; x.bytes = L.total;
  x.fs.nu = (size_t)(nx + 1) * ny; x.fs.nv = (size_t)nx * (ny + 1);
  x.fs.nup = padded_faces(x.fs.nu); x.fs.nvp = padded_faces(x.fs.nv);
  x.fs.u = base + L.u; x.fs.v = base + L.v;
  x.fs.mono_u = (uint8_t*)(base + L.mono_u); x.fs.mono_v = (uint8_t*)(base + L.mono_v);
  x.fs.khtr_u = (double*)(base + L.ku); x.fs.khtr_v = (double*)(base + L.kv);
  x.ws = base + L.ws;
  rc = fill_face_state_form(x.fs, nz, b->form, (hipStream_t)stream);
  if (rc) {
    b->arena.give(off, L.total);
    x = BatchDomain{};
    return rc;
  }
  x.live = true;
  ++b->ndom;
  b->dirty = true;
  return slot;
}

int redi_batch_remove(redi_batch_t* b, int slot) {
  if (!b || slot < 0 || slot >= b->max_domains || !b->dom[slot].live)
    return fail(REDI_ERR_ARG, "redi_batch_remove: no domain in slot %d", slot);
  BatchDomain& x = b->dom[slot];
  b->arena.give(x.off, x.bytes);
  x = BatchDomain{};
  --b->ndom;
  b->dirty = true;
  return REDI_OK;
}

int redi_batch_step(redi_batch_t* b, int nsteps, void* stream) {
  if (!b) return fail(REDI_ERR_ARG, "redi_batch_step: null handle");
  const Range range("redi_batch_step");
  if (nsteps < 0) return fail(REDI_ERR_ARG, "redi_batch_step: nsteps %d < 0", nsteps);
  hipStream_t st = (hipStream_t)stream;
  if (b->dirty) {
    const int rc = batch_upload(b, st);
    if (rc) return rc;
  }
  if (b->ndom == 0 || nsteps == 0) return REDI_OK;
  if (b->graph_stale || !b->exec) {
    const int rc = with_batch_ops(b->form, b->group, [&](auto ops) { return decltype(ops)::capture(b); });
    if (rc) return rc;
  }
  for (int n = 0; n < nsteps; ++n) HIP_TRY(hipGraphLaunch(b->exec, st));
  return REDI_OK;
}

int redi_batch_step_serial(redi_batch_t* b, int nsteps, void* stream) {
  if (!b) return fail(REDI_ERR_ARG, "redi_batch_step_serial: null handle");
  const Range range("redi_batch_step_serial");
  if (nsteps < 0) return fail(REDI_ERR_ARG, "redi_batch_step_serial: nsteps %d < 0", nsteps);
  for (int n = 0; n < nsteps; ++n) {
    const int rc = with_batch_ops(b->form, b->group,
                                  [&](auto ops) { return decltype(ops)::serial(b, (hipStream_t)stream); });
    if (rc) return rc;
  }
  HIP_TRY(hipGetLastError());
  return REDI_OK;
}

int redi_batch_export_coeffs(const redi_batch_t* b, int slot, const redi_coeffs_t* dst, void* stream) {
  if (!b || !dst || slot < 0 || slot >= b->max_domains || !b->dom[slot].live)
    return fail(REDI_ERR_ARG, "redi_batch_export_coeffs: no domain in slot %d", slot);
  if (!dst->uPoL || !dst->uPoR || !dst->uKoL || !dst->uKoR || !dst->uhEff || !dst->vPoL ||
      !dst->vPoR || !dst->vKoL || !dst->vKoR || !dst->vhEff)
    return fail(REDI_ERR_ARG, "redi_batch_export_coeffs: incomplete destination arrays");
  const BatchDomain& x = b->dom[slot];
  if (b->form == REDI_FORM_FP32_STORE)
    return export_coeffs<Fp32>(x.fs, x.nx, x.ny, x.nz, *dst, true, true, (hipStream_t)stream);
  return export_coeffs<Fp64>(x.fs, x.nx, x.ny, x.nz, *dst, true, true, (hipStream_t)stream);
}

int redi_batch_info(const redi_batch_t* b, int* ndomains, long long* used, long long* largest) {
  if (!b) return fail(REDI_ERR_ARG, "redi_batch_info: null handle");
  if (ndomains) *ndomains = b->ndom;
  if (used) *used = (long long)b->arena.used();
  if (largest) *largest = (long long)b->arena.largest();
  return REDI_OK;
}

}  // extern "C"
