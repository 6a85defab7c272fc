// simt.cpp -- fiber scheduler and runtime API of the host test build (see
// simt.hpp for the execution model).  Test infrastructure only.
#include "simt.hpp"

#include <setjmp.h>
#include <stdarg.h>
#include <ucontext.h>

#include <chrono>
#include <numeric>
#include <random>

#include <unordered_map>

#if defined(__SANITIZE_ADDRESS__)
#include <sanitizer/common_interface_defs.h>
#define SIMT_ASAN 1
#else
#define SIMT_ASAN 0
#endif
#if defined(__has_feature)
#if __has_feature(memory_sanitizer)
#include <sanitizer/msan_interface.h>
#define SIMT_MSAN 1
#endif
#endif
#ifndef SIMT_MSAN
#define SIMT_MSAN 0
#endif

namespace simt {

dim3 g_thread, g_block, g_blockdim, g_griddim;

namespace {

constexpr size_t kStackBytes = 512 * 1024;
enum State { kReady, kAtBarrier, kAtWaveOp, kDone };

// A fiber is entered the first time with setcontext (its own stack and
// entry point); every later switch is a _setjmp / _longjmp pair, which
// leaves the signal mask alone (swapcontext makes a system call per switch).
struct Fiber {
  ucontext_t ctx;
  jmp_buf jb;
  bool started = false;
  State st = kReady;
  int line = 0;
  void* fake = nullptr;  // the sanitizer's fake-stack handle while switched out
};

struct WaveSlot {
  int arrived = 0;
  int kind = 0, line = 0;
  size_t bytes = 0;
  bool here[64];
  int arg[64];
  uint8_t in[64][16];
  uint8_t out[64][16];
};

struct Group {
  std::vector<Fiber> f;
  int n = 0, live = 0;
  int bar_arrived = 0, bar_line = -1;
  std::vector<WaveSlot> w;
  std::vector<int> wave_live;
  const std::function<void()>* body = nullptr;
};

jmp_buf g_sched;
Group g_grp;
int g_cur = -1;
int g_policy = 0;
std::mt19937_64 g_rng(1);
long long g_inactive = 0, g_syncs = 0;
uint64_t g_block_gen = 0;  // counts workgroups started
std::unordered_map<void*, uint64_t> g_lds_gen;
struct Stacks {  // fiber stacks, reused by every workgroup
  std::vector<char*> v;
  ~Stacks() {
    for (char* p : v) std::free(p);
  }
} g_stack_pool;
std::vector<char*>& g_stacks = g_stack_pool.v;
std::vector<uint8_t> g_dyn;
const void* g_sched_bottom = nullptr;
size_t g_sched_size = 0;
hipError_t g_last = hipSuccess;

[[noreturn]] void fatal(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  std::fprintf(stderr, "simt: ");
  std::vfprintf(stderr, fmt, ap);
  std::fprintf(stderr, " (block %u of %u, %u threads)\n", g_block.x, g_griddim.x, g_blockdim.x);
  va_end(ap);
  std::fflush(stderr);
  std::_Exit(3);
}

void to_fiber(int t) {
  Fiber& f = g_grp.f[t];
  g_cur = t;
  g_thread = dim3((unsigned)t, 0, 0);
#if SIMT_ASAN
  void* fake = nullptr;
  __sanitizer_start_switch_fiber(&fake, g_stacks[t], kStackBytes);
#endif
  if (!_setjmp(g_sched)) {
    if (!f.started) {
      f.started = true;
      setcontext(&f.ctx);
    }
    _longjmp(f.jb, 1);
  }
#if SIMT_ASAN
  __sanitizer_finish_switch_fiber(fake, nullptr, nullptr);
#endif
  g_cur = -1;
}

void to_sched(bool dying) {
  Fiber& f = g_grp.f[g_cur];
#if SIMT_ASAN
  __sanitizer_start_switch_fiber(dying ? nullptr : &f.fake, g_sched_bottom, g_sched_size);
#else
  (void)dying;
#endif
  if (!_setjmp(f.jb)) _longjmp(g_sched, 1);
#if SIMT_ASAN
  __sanitizer_finish_switch_fiber(f.fake, nullptr, nullptr);
#endif
}

void release_barrier() {
  Group& g = g_grp;
  if (g.bar_arrived == 0 || g.bar_arrived != g.live) return;
  for (auto& f : g.f)
    if (f.st == kAtBarrier) f.st = kReady;
  g.bar_arrived = 0;
  g.bar_line = -1;
  ++g_syncs;
}

void release_wave(int w) {
  Group& g = g_grp;
  WaveSlot& s = g.w[(size_t)w];
  if (s.arrived == 0 || s.arrived != g.wave_live[(size_t)w]) return;
  const int base = w * 64;
  const int lanes = std::min(64, g.n - base);
  uint64_t mask = 0;
  int first = -1;
  for (int l = 0; l < lanes; ++l)
    if (s.here[l]) {
      if (first < 0) first = l;
      if (s.kind == kOpBallot && s.in[l][0]) mask |= 1ull << l;
    }
  for (int l = 0; l < lanes; ++l) {
    if (!s.here[l]) continue;
    int src = l;
    switch (s.kind) {
      case kOpBallot:
        std::memcpy(s.out[l], &mask, 8);
        continue;
      case kOpShfl: src = s.arg[l] & 63; break;
      case kOpShflDown: src = (l + s.arg[l] >= 64) ? l : l + s.arg[l]; break;
      case kOpReadFirst: src = first; break;
      default: fatal("unknown cross-lane operation %d", s.kind);
    }
    if (src < lanes && s.here[src]) {
      std::memcpy(s.out[l], s.in[src], s.bytes);
    } else {
      std::memset(s.out[l], 0xFF, s.bytes);
      ++g_inactive;
    }
  }
  for (int l = 0; l < lanes; ++l)
    if (s.here[l]) g.f[(size_t)(base + l)].st = kReady;
  s.arrived = 0;
  ++g_syncs;
}

void trampoline() {
#if SIMT_ASAN
  __sanitizer_finish_switch_fiber(nullptr, &g_sched_bottom, &g_sched_size);
#endif
  (*g_grp.body)();
  Group& g = g_grp;
  g.f[(size_t)g_cur].st = kDone;
  --g.live;
  --g.wave_live[(size_t)(g_cur / 64)];
  release_barrier();
  release_wave(g_cur / 64);
  to_sched(true);
  fatal("a returned thread was resumed");
}

[[noreturn]] void deadlock() {
  std::fprintf(stderr, "simt: workgroup cannot make progress:\n");
  for (size_t t = 0; t < g_grp.f.size(); ++t) {
    const Fiber& f = g_grp.f[t];
    if (f.st == kAtBarrier || f.st == kAtWaveOp)
      std::fprintf(stderr, "  thread %zu waits at %s, line %d\n", t,
                   f.st == kAtBarrier ? "a barrier" : "a cross-lane operation", f.line);
  }
  fatal("deadlock");
}

}  // namespace

void set_policy(int policy, uint64_t seed) {
  g_policy = policy;
  g_rng.seed(seed);
}
long long inactive_reads() { return g_inactive; }
long long sync_points() { return g_syncs; }

void* dynamic_lds() { return g_dyn.empty() ? nullptr : g_dyn.data(); }

void undefined_bytes(void* p, size_t n) {
  if (!n) return;
  std::memset(p, 0x7F, n);
#if SIMT_MSAN
  __msan_poison(p, n);
#endif
}

void lds_region(void* p, size_t n) {
  uint64_t& g = g_lds_gen[p];
  if (g == g_block_gen) return;
  g = g_block_gen;
  undefined_bytes(p, n);
}

void barrier(int line) {
  if (g_cur < 0) fatal("__syncthreads outside a kernel");
  Group& g = g_grp;
  Fiber& f = g.f[(size_t)g_cur];
  if (g.bar_arrived == 0)
    g.bar_line = line;
  else if (line != g.bar_line)
    fatal("divergent barrier: thread %d at line %d, others at line %d", g_cur, line, g.bar_line);
  f.st = kAtBarrier;
  f.line = line;
  ++g.bar_arrived;
  release_barrier();
  to_sched(false);
}

void wave_op(int kind, int line, const void* in, size_t bytes, int arg, void* out) {
  if (g_cur < 0) fatal("cross-lane operation outside a kernel");
  Group& g = g_grp;
  const int t = g_cur, w = t / 64, lane = t % 64;
  WaveSlot& s = g.w[(size_t)w];
  if (s.arrived == 0) {
    s.kind = kind;
    s.line = line;
    s.bytes = bytes;
    std::fill(std::begin(s.here), std::end(s.here), false);
  } else if (s.kind != kind || s.line != line) {
    fatal("divergent cross-lane operation: lane %d at line %d, others at line %d", lane, line,
          s.line);
  }
  std::memcpy(s.in[lane], in, bytes);
  s.arg[lane] = arg;
  s.here[lane] = true;
  ++s.arrived;
  Fiber& f = g.f[(size_t)t];
  f.st = kAtWaveOp;
  f.line = line;
  release_wave(w);
  to_sched(false);
  std::memcpy(out, s.out[lane], kind == kOpBallot ? 8 : bytes);
}

void run_grid(dim3 grid, dim3 block, size_t dyn_lds, const std::function<void()>& body) {
  if (grid.y != 1 || grid.z != 1 || block.y != 1 || block.z != 1)
    fatal("multi-dimensional launch (%u,%u,%u) x (%u,%u,%u)", grid.x, grid.y, grid.z, block.x,
          block.y, block.z);
  const int n = (int)block.x;
  if (n == 0 || grid.x == 0) return;
  if (n > 1024) fatal("workgroup of %d threads", n);
  while ((int)g_stacks.size() < n) g_stacks.push_back((char*)std::malloc(kStackBytes));
  g_dyn.assign(dyn_lds, 0);
  std::vector<unsigned> order(grid.x);
  std::iota(order.begin(), order.end(), 0u);
  if (g_policy == 1) std::reverse(order.begin(), order.end());
  if (g_policy == 2) std::shuffle(order.begin(), order.end(), g_rng);
  g_griddim = grid;
  g_blockdim = block;
  Group& g = g_grp;
  std::vector<int> ready;
  ready.reserve((size_t)n);
  for (unsigned b : order) {
    g_block = dim3(b, 0, 0);
    ++g_block_gen;
    if (!g_dyn.empty()) undefined_bytes(g_dyn.data(), g_dyn.size());
    g.n = n;
    g.live = n;
    g.f.assign((size_t)n, Fiber{});
    const int nw = (n + 63) / 64;
    g.w.assign((size_t)nw, WaveSlot{});
    g.wave_live.assign((size_t)nw, 0);
    for (int w = 0; w < nw; ++w) g.wave_live[(size_t)w] = std::min(64, n - 64 * w);
    g.bar_arrived = 0;
    g.bar_line = -1;
    g.body = &body;
    for (int t = 0; t < n; ++t) {
      Fiber& f = g.f[(size_t)t];
      getcontext(&f.ctx);
      f.ctx.uc_stack.ss_sp = g_stacks[(size_t)t];
      f.ctx.uc_stack.ss_size = kStackBytes;
      f.ctx.uc_link = nullptr;
      makecontext(&f.ctx, trampoline, 0);
      f.st = kReady;
    }
    for (;;) {
      ready.clear();
      for (int t = 0; t < n; ++t)
        if (g.f[(size_t)t].st == kReady) ready.push_back(t);
      if (ready.empty()) {
        if (g.live == 0) break;
        deadlock();
      }
      int pick = ready.front();
      if (g_policy == 1) pick = ready.back();
      if (g_policy == 2) pick = ready[(size_t)(g_rng() % ready.size())];
      to_fiber(pick);
    }
  }
  g_thread = dim3(0, 0, 0);
}

// ---------------------------------------------------------------------------
// memory, streams, events, graphs
// ---------------------------------------------------------------------------
void* dev_alloc(size_t bytes) {
  void* p = nullptr;
  if (posix_memalign(&p, 256, bytes ? bytes : 1) != 0) return nullptr;
  return p;
}
void dev_free(void* p) { std::free(p); }

}  // namespace simt

struct simt_stream {
  simt_graph* cap = nullptr;  // non-null while captured
};
struct simt_event {
  double ms = 0.0;
};
struct simt_graph {
  std::vector<std::function<void()>> nodes;
};

namespace {
simt_stream g_null_stream;
simt_stream* stream_of(hipStream_t st) { return st ? st : &g_null_stream; }
double now_ms() {
  using namespace std::chrono;
  return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}
}  // namespace

void simt::submit(hipStream_t st, std::function<void()> fn) {
  simt_stream* s = stream_of(st);
  if (s->cap)
    s->cap->nodes.push_back(std::move(fn));
  else
    fn();
}

hipError_t hipFree(void* p) {
  simt::dev_free(p);
  return hipSuccess;
}
hipError_t hipHostMalloc(void** p, size_t bytes, unsigned) {
  *p = simt::dev_alloc(bytes);
  return *p ? hipSuccess : hipErrorOutOfMemory;
}
hipError_t hipHostFree(void* p) {
  simt::dev_free(p);
  return hipSuccess;
}
hipError_t hipMemcpy(void* dst, const void* src, size_t bytes, hipMemcpyKind) {
  if (bytes) std::memmove(dst, src, bytes);
  return hipSuccess;
}
hipError_t hipMemcpyAsync(void* dst, const void* src, size_t bytes, hipMemcpyKind,
                          hipStream_t st) {
  simt::submit(st, [=]() {
    if (bytes) std::memmove(dst, src, bytes);
  });
  return hipSuccess;
}
hipError_t hipMemset(void* dst, int v, size_t bytes) {
  if (bytes) std::memset(dst, v, bytes);
  return hipSuccess;
}
hipError_t hipMemsetAsync(void* dst, int v, size_t bytes, hipStream_t st) {
  simt::submit(st, [=]() {
    if (bytes) std::memset(dst, v, bytes);
  });
  return hipSuccess;
}
hipError_t hipGetLastError() {
  const hipError_t e = simt::g_last;
  simt::g_last = hipSuccess;
  return e;
}
hipError_t hipPeekAtLastError() { return simt::g_last; }
const char* hipGetErrorString(hipError_t e) {
  switch (e) {
    case hipSuccess: return "hipSuccess";
    case hipErrorInvalidValue: return "hipErrorInvalidValue";
    case hipErrorOutOfMemory: return "hipErrorOutOfMemory";
    case hipErrorNotSupported: return "hipErrorNotSupported";
    default: return "hipErrorUnknown";
  }
}
hipError_t hipGetDevice(int* d) {
  *d = 0;
  return hipSuccess;
}
hipError_t hipSetDevice(int d) { return d == 0 ? hipSuccess : hipErrorInvalidValue; }
hipError_t hipGetDeviceProperties(hipDeviceProp_t* p, int d) {
  if (d != 0) return hipErrorInvalidValue;
  std::memset(p, 0, sizeof(*p));
  std::snprintf(p->name, sizeof(p->name), "simt host build");
  p->totalGlobalMem = (size_t)8 << 30;
  p->sharedMemPerBlock = 160 * 1024;
  p->warpSize = 64;
  p->maxThreadsPerBlock = 1024;
  p->multiProcessorCount = 4;  // small persistent grids: every item strides
  p->l2CacheSize = 4 << 20;
  p->clockRate = 2400000;
  return hipSuccess;
}
hipError_t hipDeviceSynchronize() { return hipSuccess; }
hipError_t hipStreamCreate(hipStream_t* st) {
  *st = new simt_stream;
  return hipSuccess;
}
hipError_t hipStreamCreateWithFlags(hipStream_t* st, unsigned) { return hipStreamCreate(st); }
hipError_t hipStreamDestroy(hipStream_t st) {
  delete st;
  return hipSuccess;
}
hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
hipError_t hipEventCreate(hipEvent_t* e) {
  *e = new simt_event;
  return hipSuccess;
}
hipError_t hipEventCreateWithFlags(hipEvent_t* e, unsigned) { return hipEventCreate(e); }
hipError_t hipEventDestroy(hipEvent_t e) {
  delete e;
  return hipSuccess;
}
hipError_t hipEventRecord(hipEvent_t e, hipStream_t st) {
  simt::submit(st, [e]() { e->ms = now_ms(); });
  return hipSuccess;
}
hipError_t hipEventSynchronize(hipEvent_t) { return hipSuccess; }
hipError_t hipEventElapsedTime(float* ms, hipEvent_t a, hipEvent_t b) {
  *ms = (float)(b->ms - a->ms);
  return hipSuccess;
}
hipError_t hipStreamWaitEvent(hipStream_t, hipEvent_t, unsigned) { return hipSuccess; }
hipError_t hipStreamBeginCapture(hipStream_t st, hipStreamCaptureMode) {
  simt_stream* s = stream_of(st);
  if (s->cap) return hipErrorInvalidValue;
  s->cap = new simt_graph;
  return hipSuccess;
}
hipError_t hipStreamEndCapture(hipStream_t st, hipGraph_t* g) {
  simt_stream* s = stream_of(st);
  *g = s->cap;
  s->cap = nullptr;
  return *g ? hipSuccess : hipErrorInvalidValue;
}
hipError_t hipGraphInstantiate(hipGraphExec_t* exec, hipGraph_t g, hipGraphNode_t*, char*, size_t) {
  *exec = new simt_graph(*g);
  return hipSuccess;
}
hipError_t hipGraphLaunch(hipGraphExec_t exec, hipStream_t st) {
  for (auto& fn : exec->nodes) simt::submit(st, fn);
  return hipSuccess;
}
hipError_t hipGraphDestroy(hipGraph_t g) {
  delete g;
  return hipSuccess;
}
hipError_t hipGraphExecDestroy(hipGraphExec_t exec) {
  delete exec;
  return hipSuccess;
}
hipError_t hipGetSymbolAddress(void** p, const void* symbol) {
  *p = const_cast<void*>(symbol);
  return hipSuccess;
}
