// simt.hpp -- runs the library's HIP kernels on the host, for CPU tests.
//
// Test infrastructure only: the product is built with hipcc for gfx950 and
// never includes this file.  A CPU test build (tests/simt/build_simt.py)
// compiles the library's own sources with g++ against this header (through
// tests/simt/hip/hip_runtime.h), under AddressSanitizer and
// UndefinedBehaviorSanitizer.
//
// Execution model.  A launch runs its workgroups one after another, in an
// order the test picks.  The threads of a workgroup run as fibers on one OS
// thread; wavefronts are 64 consecutive threads.  A fiber runs until it
// reaches a synchronisation point -- __syncthreads, a cross-lane operation
// (__ballot, __shfl, __shfl_down, readfirstlane) -- or returns.  A barrier
// releases when every thread of the workgroup that has not returned has
// reached it; a cross-lane operation releases when every lane of the
// wavefront that has not returned has reached it, and its results are
// computed at that moment from the values every lane brought.  Which fiber
// runs next is chosen by the policy (lowest thread first, highest first, or
// seeded random), so any exchange of data between lanes that is not ordered
// by one of those points -- LDS written by one lane and read by another in
// the same stretch, global memory shared between workgroups -- gives results
// that depend on the policy, and tests compare the policies.
//
// Checked and reported (the process exits with status 3):
//   * a barrier or cross-lane operation reached by the threads of one
//     workgroup / wavefront at different source lines (divergent control
//     flow around it);
//   * a workgroup that cannot make progress (some threads wait at a barrier
//     that the others never reach);
//   * kernels with a multi-dimensional grid or workgroup (not used here).
//   * under MemorySanitizer (msan builds): a branch, address or output that
//     depends on LDS or device memory no thread wrote.
// Counted and reported: __shfl / __shfl_down reading a lane that has
// returned (the GPU returns that lane's stale register; here the result is
// poisoned with 0xFF bytes, NaN for floating point).
#pragma once

#include <math.h>
#include <stddef.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <cfenv>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <cstdlib>
#include <functional>
#include <memory>
#include <tuple>
#include <type_traits>
#include <utility>
#include <vector>

// ---------------------------------------------------------------------------
// Language keywords
// ---------------------------------------------------------------------------
#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __launch_bounds__(...)
#define __align__(n) __attribute__((aligned(n)))
// One workgroup runs at a time, so a function-static object is that
// workgroup's LDS.  The host build also follows every __shared__
// declaration with simt::lds_region(&x, sizeof x), which gives the object
// undefined contents at the start of each workgroup, as LDS has.
#define __shared__ static
#define HIP_SYMBOL(x) ((const void*)&(x))

struct dim3 {
  unsigned x, y, z;
  constexpr dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {}
};
struct uint4 {
  unsigned x, y, z, w;
};
struct uint2 {
  unsigned x, y;
};
struct int2 {
  int x, y;
};
struct float2 {
  float x, y;
};
struct double2 {
  double x, y;
};

namespace simt {

extern dim3 g_thread, g_block, g_blockdim, g_griddim;

// Scheduling policy of lanes and workgroups: 0 lowest first, 1 highest
// first, 2 seeded random.
void set_policy(int policy, uint64_t seed);
// Shuffles that read a returned lane, since the start of the process.
long long inactive_reads();
// Synchronisation points released, since the start of the process.
long long sync_points();

// Cross-lane operations.  `in` / `out` are the calling lane's bytes.
enum OpKind { kOpBallot = 1, kOpShfl, kOpShflDown, kOpReadFirst };
void wave_op(int kind, int line, const void* in, size_t bytes, int arg, void* out);
void barrier(int line);
void* dynamic_lds();
// The first thread of a workgroup to declare LDS object [p, p + n) fills it
// with 0x7F bytes (under MemorySanitizer: marks it uninitialized).
void lds_region(void* p, size_t n);

inline uint64_t ballot(bool pred, int line) {
  uint64_t r = 0;
  const uint8_t p = pred ? 1 : 0;
  wave_op(kOpBallot, line, &p, 1, 0, &r);
  return r;
}
template <class T>
inline T shfl(T v, int src, int line) {
  static_assert(sizeof(T) <= 16, "shuffled value too large");
  T r;
  wave_op(kOpShfl, line, &v, sizeof(T), src, &r);
  return r;
}
template <class T>
inline T shfl_down(T v, unsigned delta, int line) {
  static_assert(sizeof(T) <= 16, "shuffled value too large");
  T r;
  wave_op(kOpShflDown, line, &v, sizeof(T), (int)delta, &r);
  return r;
}
template <class T>
inline T readfirstlane(T v, int line) {
  T r;
  wave_op(kOpReadFirst, line, &v, sizeof(T), 0, &r);
  return r;
}
inline unsigned lane_id() { return g_thread.x & 63u; }
// bits of m below this lane (lanes 0..31: mask bits of lanes 0..31)
inline unsigned mbcnt_lo(unsigned m, unsigned acc) {
  const unsigned l = lane_id();
  const unsigned below = l >= 32 ? m : (m & ((1u << l) - 1u));
  return acc + (unsigned)__builtin_popcount(below);
}
// (lanes 32..63: mask bits of lanes 32..63)
inline unsigned mbcnt_hi(unsigned m, unsigned acc) {
  const unsigned l = lane_id();
  if (l < 32) return acc;
  const unsigned below = m & ((1u << (l - 32)) - 1u);
  return acc + (unsigned)__builtin_popcount(below);
}

// Runs `body` for every thread of a 1-D grid of 1-D workgroups.
void run_grid(dim3 grid, dim3 block, size_t dyn_lds, const std::function<void()>& body);

}  // namespace simt

#define threadIdx (::simt::g_thread)
#define blockIdx (::simt::g_block)
#define blockDim (::simt::g_blockdim)
#define gridDim (::simt::g_griddim)

#define __syncthreads() ::simt::barrier(__LINE__)
#define __ballot(p) ::simt::ballot((p), __LINE__)
#define __shfl(v, src, ...) ::simt::shfl((v), (src), __LINE__)
#define __shfl_down(v, d, ...) ::simt::shfl_down((v), (unsigned)(d), __LINE__)
#define __builtin_amdgcn_readfirstlane(v) ::simt::readfirstlane((v), __LINE__)
#define __builtin_amdgcn_mbcnt_lo(m, a) ::simt::mbcnt_lo((m), (a))
#define __builtin_amdgcn_mbcnt_hi(m, a) ::simt::mbcnt_hi((m), (a))
// the hardware reciprocal (within 1 ulp) as the quotient
#define __builtin_amdgcn_rcpf(x) (1.0f / (x))
// a wavefront-scope fence: the lanes of a wave exchange data through LDS
// across it, which here needs the fibers to meet (a barrier of the
// workgroup, stronger than the wave's in-order execution it stands for)
#define __builtin_amdgcn_fence(order, scope) ::simt::barrier(__LINE__)
#define __popcll(x) __builtin_popcountll(x)
// conversion rounding to nearest, ties to even (the default rounding mode)
inline unsigned __double2uint_rn(double x) { return (unsigned)std::nearbyint(x); }
#define __popc(x) __builtin_popcount(x)

// ---------------------------------------------------------------------------
// Runtime API (the subset the library calls); device memory is host memory.
// ---------------------------------------------------------------------------
typedef enum hipError_t {
  hipSuccess = 0,
  hipErrorInvalidValue = 1,
  hipErrorOutOfMemory = 2,
  hipErrorNotSupported = 801,
  hipErrorUnknown = 999
} hipError_t;
typedef enum hipMemcpyKind {
  hipMemcpyHostToHost = 0,
  hipMemcpyHostToDevice = 1,
  hipMemcpyDeviceToHost = 2,
  hipMemcpyDeviceToDevice = 3,
  hipMemcpyDefault = 4
} hipMemcpyKind;
typedef enum hipStreamCaptureMode {
  hipStreamCaptureModeGlobal = 0,
  hipStreamCaptureModeThreadLocal = 1,
  hipStreamCaptureModeRelaxed = 2
} hipStreamCaptureMode;
#define hipStreamNonBlocking 0x01
#define hipEventDisableTiming 0x02
#define hipHostMallocDefault 0x0

struct simt_stream;
struct simt_event;
struct simt_graph;
struct simt_graph_node;
typedef simt_stream* hipStream_t;
typedef simt_event* hipEvent_t;
typedef simt_graph* hipGraph_t;
typedef simt_graph* hipGraphExec_t;
typedef simt_graph_node* hipGraphNode_t;

struct hipDeviceProp_t {
  char name[256];
  size_t totalGlobalMem;
  size_t sharedMemPerBlock;
  int warpSize;
  int maxThreadsPerBlock;
  int multiProcessorCount;
  int l2CacheSize;
  int clockRate;
};

namespace simt {
void* dev_alloc(size_t bytes);
void dev_free(void* p);
// Runs `fn` now, or records it when `st` is being captured.
void submit(hipStream_t st, std::function<void()> fn);
}  // namespace simt

template <class T>
inline hipError_t hipMalloc(T** p, size_t bytes) {
  *p = (T*)simt::dev_alloc(bytes);
  return (*p || bytes == 0) ? hipSuccess : hipErrorOutOfMemory;
}
hipError_t hipFree(void* p);
hipError_t hipHostMalloc(void** p, size_t bytes, unsigned flags);
hipError_t hipHostFree(void* p);
hipError_t hipMemcpy(void* dst, const void* src, size_t bytes, hipMemcpyKind kind);
hipError_t hipMemcpyAsync(void* dst, const void* src, size_t bytes, hipMemcpyKind kind,
                          hipStream_t st);
hipError_t hipMemset(void* dst, int v, size_t bytes);
hipError_t hipMemsetAsync(void* dst, int v, size_t bytes, hipStream_t st);
hipError_t hipGetLastError();
hipError_t hipPeekAtLastError();
const char* hipGetErrorString(hipError_t e);
hipError_t hipGetDevice(int* d);
hipError_t hipSetDevice(int d);
hipError_t hipGetDeviceProperties(hipDeviceProp_t* p, int d);
hipError_t hipDeviceSynchronize();
hipError_t hipStreamCreate(hipStream_t* st);
hipError_t hipStreamCreateWithFlags(hipStream_t* st, unsigned flags);
hipError_t hipStreamDestroy(hipStream_t st);
hipError_t hipStreamSynchronize(hipStream_t st);
hipError_t hipEventCreate(hipEvent_t* e);
hipError_t hipEventCreateWithFlags(hipEvent_t* e, unsigned flags);
hipError_t hipEventDestroy(hipEvent_t e);
hipError_t hipEventRecord(hipEvent_t e, hipStream_t st);
hipError_t hipEventSynchronize(hipEvent_t e);
hipError_t hipEventElapsedTime(float* ms, hipEvent_t a, hipEvent_t b);
hipError_t hipStreamWaitEvent(hipStream_t st, hipEvent_t e, unsigned flags);
hipError_t hipStreamBeginCapture(hipStream_t st, hipStreamCaptureMode mode);
hipError_t hipStreamEndCapture(hipStream_t st, hipGraph_t* g);
hipError_t hipGraphInstantiate(hipGraphExec_t* exec, hipGraph_t g, hipGraphNode_t* err_node,
                               char* log, size_t log_size);
hipError_t hipGraphLaunch(hipGraphExec_t exec, hipStream_t st);
hipError_t hipGraphDestroy(hipGraph_t g);
hipError_t hipGraphExecDestroy(hipGraphExec_t exec);
hipError_t hipGetSymbolAddress(void** p, const void* symbol);

template <class K>
inline hipError_t hipOccupancyMaxActiveBlocksPerMultiprocessor(int* n, K, int, size_t) {
  *n = 2;
  return hipSuccess;
}

// ---------------------------------------------------------------------------
// Kernel launches: the build rewrites `k<<<g, b, lds, st>>>(args)` into
// `simt::launch(k, g, b, lds, st)(args)`.
// ---------------------------------------------------------------------------
namespace simt {
template <class... P>
struct Launch {
  void (*kernel)(P...);
  dim3 grid, block;
  size_t lds;
  hipStream_t st;
  template <class... A>
  void operator()(A&&... a) const {
    auto args = std::make_shared<std::tuple<std::decay_t<P>...>>(std::forward<A>(a)...);
    void (*k)(P...) = kernel;
    const dim3 g = grid, b = block;
    const size_t l = lds;
    submit(st, [k, args, g, b, l]() {
      run_grid(g, b, l, [k, args]() { std::apply(k, *args); });
    });
  }
};
template <class... P>
inline Launch<P...> launch(void (*k)(P...), dim3 grid, dim3 block, size_t lds = 0,
                           hipStream_t st = nullptr) {
  return Launch<P...>{k, grid, block, lds, st};
}
}  // namespace simt
