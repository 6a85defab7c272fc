// selftest.cpp -- small kernels that show what the host test build catches
// (tests/test_simt_cpu.py runs each mode and checks the outcome).  It is
// built through the same source rewrite as the library.
//
//   simt_selftest MODE [policy]
//
// Modes print one line of results, or end in the failure they plant:
//   race        lane l writes LDS[l], reads LDS[l+1] with no barrier between
//               (the result depends on the lane order)
//   race_sync   the same with __syncthreads between (it does not)
//   shfl        __shfl_down by 1 and __shfl from lane 63-l, FP64
//   compact     ballot + mbcnt compaction of every third lane
//   returned    lanes 32..63 return, lane 31 shuffles down from lane 32
//   divergent   half a wavefront at one __syncthreads, half at another
//   oob         LDS array of 64 written at index lane+1 (AddressSanitizer)
//   uninit      a branch on device memory no one wrote (MemorySanitizer)
//   lds_uninit  a branch on an LDS word no lane wrote (MemorySanitizer; in
//               other builds a workgroup's LDS starts as 0x7F bytes)
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "simt.hpp"

__global__ void k_race(int* out, int sync) {
  __shared__ int lds[64];
  const int l = (int)threadIdx.x;
  lds[l] = l + 1;
  if (sync) __syncthreads();
  out[l] = lds[(l + 1) % 64];
}

__global__ void k_shfl(double* out) {
  const int l = (int)threadIdx.x;
  const double v = (double)l;
  out[l] = __shfl_down(v, 1, 64);
  out[64 + l] = __shfl(v, 63 - l, 64);
}

__global__ void k_compact(int* out, unsigned long long* mask) {
  const int l = (int)threadIdx.x;
  const bool p = l % 3 == 0;
  const uint64_t b = __ballot(p);
  const int pos = (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(b >> 32),
                                                 __builtin_amdgcn_mbcnt_lo((uint32_t)b, 0u));
  if (p) out[pos] = l;
  if (l == 0) *mask = b;
}

__global__ void k_returned(double* out) {
  const int l = (int)threadIdx.x;
  if (l >= 32) return;
  out[l] = __shfl_down((double)l, 1, 64);
}

__global__ void k_divergent() {
  if (threadIdx.x < 32) {
    __syncthreads();
  } else {
    __syncthreads();
  }
}

__global__ void k_oob(int* out) {
  __shared__ int lds[64];
  const int l = (int)threadIdx.x;
  lds[l + 1] = l;
  out[l] = 0;
}

__global__ void k_uninit(const int* in, int* out) {
  const int l = (int)threadIdx.x;
  out[(unsigned)in[l] % 64u] = 1;  // an address from memory no one wrote
}

__global__ void k_lds_uninit(int* out) {
  __shared__ int lds[64];
  const int l = (int)threadIdx.x;
  if (l % 2) lds[l] = 1;
  __syncthreads();
  out[(unsigned)lds[l] % 64u] = l;  // even lanes: an address from an LDS word no lane wrote
}

int main(int argc, char** argv) {
  if (argc < 2) return 1;
  const char* m = argv[1];
  simt::set_policy(argc > 2 ? atoi(argv[2]) : 0, 11);
  if (!strcmp(m, "race") || !strcmp(m, "race_sync")) {
    int out[64];
    simt::launch(k_race, 1, 64, 0, nullptr)(out, strcmp(m, "race") ? 1 : 0);
    for (int l = 0; l < 64; ++l) printf("%d ", out[l]);
  } else if (!strcmp(m, "shfl")) {
    double out[128];
    simt::launch(k_shfl, 1, 64, 0, nullptr)(out);
    for (double v : out) printf("%g ", v);
  } else if (!strcmp(m, "compact")) {
    int out[64];
    memset(out, 0xFF, sizeof(out));
    unsigned long long mask = 0;
    simt::launch(k_compact, 1, 64, 0, nullptr)(out, &mask);
    printf("%llx ", mask);
    for (int l = 0; l < 22; ++l) printf("%d ", out[l]);
  } else if (!strcmp(m, "returned")) {
    double out[32];
    simt::launch(k_returned, 1, 64, 0, nullptr)(out);
    printf("%g %g %lld", out[30], out[31], simt::inactive_reads());
  } else if (!strcmp(m, "divergent")) {
    simt::launch(k_divergent, 1, 64, 0, nullptr)();
  } else if (!strcmp(m, "oob")) {
    int out[64];
    simt::launch(k_oob, 1, 64, 0, nullptr)(out);
  } else if (!strcmp(m, "uninit")) {
    int* in = nullptr;
    int out[64];
    if (hipMalloc(&in, 64 * sizeof(int)) != hipSuccess) return 1;
    simt::launch(k_uninit, 1, 64, 0, nullptr)(in, out);
    (void)hipFree(in);
  } else if (!strcmp(m, "lds_uninit")) {
    int out[64];
    simt::launch(k_lds_uninit, 1, 64, 0, nullptr)(out);
  } else {
    return 1;
  }
  printf("\n");
  return 0;
}
