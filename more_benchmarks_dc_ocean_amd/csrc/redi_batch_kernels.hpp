// redi_batch_kernels.hpp -- the Redi step over many independent domains of
// uneven size, one persistent launch per kernel kind.
//
// Each batched kernel reads a work table from device memory: item e (one
// domain's launch of that kernel kind -- one of its two face grids for the
// sweep and the face fluxes) owns the virtual blocks [start[e], start[e] +
// nblk[e]) of a virtual grid of count[1] blocks.  A fixed grid of workgroups
// (sized to fill the machine once) strides over the virtual blocks and runs,
// for each, the same kernel body the single-domain launch runs for that
// block, with the item's arguments.  Every result is therefore the
// single-domain result, bit for bit.
//
// The grid and the table addresses are the only launch arguments, so one
// captured hipGraph serves every batch composition: adding or retiring a
// domain rewrites the table contents, never the graph.
//
// Item starts are multiples of 8 and the persistent grid is a multiple of 8,
// so virtual block b runs on XCD b % 8 and the single-domain kernels' XCD
// banding (xcd_band over the item's own block index) keeps holding.
#pragma once

#include "redi_kernels.hpp"

namespace redi {

template <class Item>
struct BatchView {
  const uint32_t* __restrict__ count;  // [0] items, [1] virtual blocks
  const uint32_t* __restrict__ start;  // first virtual block of item e (count[0] + 1 entries)
  const uint32_t* __restrict__ nblk;   // blocks of item e (<= start[e+1] - start[e])
  const Item* __restrict__ items;
};

// The work tables do not change while a kernel runs (they are copied in
// before the launch, in stream order), so they are read through the constant
// address space: every lane needs the same entry, and scalar loads bring it
// into scalar registers once per virtual block instead of a vector load per
// field.
template <class T>
__device__ __forceinline__ T ld_table(const T* p) {
  using CT = const __attribute__((address_space(4))) T;
  return *(CT*)p;  // address-space cast (reinterpret_cast may not change it)
}

// A whole table item, word by word into scalar registers (readfirstlane
// marks every word uniform, so the body's uses never reload it).
template <class Item>
__device__ __forceinline__ Item ld_item(const Item* p) {
  static_assert(sizeof(Item) % 4 == 0, "table items are whole words");
  using CW = const __attribute__((address_space(4))) uint32_t;
  CW* w = (CW*)p;
  union {
    Item it;
    uint32_t u[sizeof(Item) / 4];
  } x;
#pragma unroll
  for (int i = 0; i < (int)(sizeof(Item) / 4); ++i) x.u[i] = __builtin_amdgcn_readfirstlane(w[i]);
  return x.it;
}

// Runs body(item, local block, item blocks) for every virtual block this
// workgroup owns.  kBarrier: the body uses LDS across waves, so the next
// block must wait until every wave is done with it.  kCopy: the item is
// copied into scalar registers first (bodies whose item is read under
// divergent branches); otherwise the body reads it in place.
template <bool kBarrier, bool kCopy = true, class Item, class Body>
__device__ __forceinline__ void for_each_block(const BatchView<Item>& v, Body&& body) {
  const uint32_t total = ld_table(v.count + 1);
  uint32_t e = 0;
  for (uint32_t b = blockIdx.x; b < total; b += gridDim.x) {
    while (ld_table(v.start + e + 1) <= b) ++e;  // b only grows, so the search only moves forward
    const uint32_t local = b - ld_table(v.start + e), nb = ld_table(v.nblk + e);
    if (local < nb) {
      if constexpr (kCopy) {
        const Item it = ld_item(v.items + e);
        body(it, (int)local, (int)nb);
      } else {
        body(v.items[e], (int)local, (int)nb);
      }
      if (kBarrier) __syncthreads();
    }
  }
}

template <class F>
struct SweepItem {
  SweepArgs<F> a;
  TileMap tm;  // tm.tx * tm.ty == kBatchSweepLanes
};
constexpr int kBatchSweepLanes = 64;

template <class F, int NT>
struct ReconItem {
  ReconArgs<F, NT> a;
  int kt;  // layer chunk: 8, 16 or 32
  int pad_;
};

template <class F>
__global__ void __launch_bounds__(128) kb_column_state(BatchView<ColumnArgs<F>> v) {
  for_each_block<false>(v, [](const ColumnArgs<F>& a, int bid, int) { column_state_body<F>(a, bid); });
}

template <class F>
__global__ void __launch_bounds__(kBatchSweepLanes) kb_neutral_sweep(BatchView<SweepItem<F>> v) {
  __shared__ uint4 stage[kSweepStageWords];
  for_each_block<false>(v, [&](const SweepItem<F>& it, int bid, int nb) {
    neutral_sweep_body<F>(it.a, it.tm, bid, nb, stage, threadIdx.x);
  });
}

__global__ void __launch_bounds__(256) kb_face_khtr(BatchView<KhtrArgs> v) {
  for_each_block<false>(v, [](const KhtrArgs& a, int bid, int) { face_khtr_body(a, bid); });
}

template <class F, int NT>
__global__ void __launch_bounds__(kReconLanes) kb_tracer_recon(BatchView<ReconItem<F, NT>> v) {
  __shared__ double lds[recon_lds_max<F, NT>()];
  for_each_block<true>(v, [&](const ReconItem<F, NT>& it, int bid, int nb) {
    switch (it.kt) {
      case 8: tracer_recon_tiled_body<F, NT, 8>(it.a, bid, nb, lds); break;
      case 32: tracer_recon_tiled_body<F, NT, 32>(it.a, bid, nb, lds); break;
      default: tracer_recon_tiled_body<F, NT, 16>(it.a, bid, nb, lds); break;
    }
  });
}

// One wavefront per workgroup (the LDS lists are the wavefront's own), so
// no barrier between blocks.
template <class F, int NT, int NZC, bool kY>
__global__ void __launch_bounds__(64) kb_flux(BatchView<FluxArgs<F>> v) {
  __shared__ FluxLds<typename F::C, NT, NZC> lds;
  for_each_block<false>(v, [&](const FluxArgs<F>& a, int bid, int nb) {
    flux_body<F, NT, NZC, kY>(a, bid, nb, lds);
  });
}

template <class F>
__global__ void __launch_bounds__(kUpdLanes) kb_tracer_update(BatchView<UpdateArgs<F>> v) {
  __shared__ double tile[kUpdTileDoubles];
  __shared__ double inv[kUpdCells];
  for_each_block<true>(v, [&](const UpdateArgs<F>& a, int bid, int nb) {
    tracer_update_body<F>(a, bid, nb, tile, inv);
  });
}

}  // namespace redi
