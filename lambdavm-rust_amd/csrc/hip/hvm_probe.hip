// hvm_probe.hip — measurement of the request rooflines of the evaluator: the
// rates at which the device performs device-scope exchanges, stores and loads
// on random slots of a net-sized buffer, the three kinds of memory-side
// request the evaluator's takes, creations, frees and links consist of.  Used
// by bench.py to report the evaluator's request rate against these ceilings
// next to the HBM roofline.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../../include/hvm_hip.h"
#include "hvm_wave.hpp"

namespace {

__device__ __forceinline__ uint32_t xorshift32(uint32_t x) {
  x ^= x << 13;
  x ^= x >> 17;
  x ^= x << 5;
  return x;
}

// Every lane keeps HVM_PROBE_INFLIGHT exchanges in flight per round, on
// pseudo-random slots of buf[0, mask]; the results are folded into a value the
// kernel writes only if it hits an impossible pattern, so no exchange is dead.
constexpr int HVM_PROBE_INFLIGHT = 8;

// OP 0: exchanges, 1: stores, 2: loads (all device-scope, relaxed).
template <typename T, int OP>
__global__ __launch_bounds__(256) void atomic_probe_kernel(T* buf, uint32_t mask, uint32_t rounds, T* sink) {
  uint32_t x = (blockIdx.x * blockDim.x + threadIdx.x) * 2654435761u + 0x9E3779B9u;
  x = x ? x : 1u;
  T acc = 0;
  for (uint32_t r = 0; r < rounds; ++r) {
    T v[HVM_PROBE_INFLIGHT];
#pragma unroll
    for (int k = 0; k < HVM_PROBE_INFLIGHT; ++k) {
      x = xorshift32(x);
      if (OP == 0) v[k] = __hip_atomic_exchange(&buf[x & mask], (T)x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      else if (OP == 1) { __hip_atomic_store(&buf[x & mask], (T)x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); v[k] = 0; }
      else v[k] = __hip_atomic_load(&buf[x & mask], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
#pragma unroll
    for (int k = 0; k < HVM_PROBE_INFLIGHT; ++k) acc ^= v[k];
  }
  if (acc == (T)0x5A5A5A5A3C3C3C3Cull) sink[0] = acc;
}

template <typename T, int OP>
int run_probe(uint64_t slots, double* ops_per_s) {
  if (slots < 1024 || (slots & (slots - 1)) || slots > (1ull << 31)) return HVM_HIP_E_ARG;
  T* buf = nullptr;
  T* sink = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = HVM_HIP_E_DEVICE;
  const uint32_t threads = 256, blocks = 8192, rounds = 16;
  if (hipMalloc(&buf, slots * sizeof(T)) != hipSuccess) return HVM_HIP_E_DEVICE;
  if (hipMalloc(&sink, sizeof(T)) == hipSuccess && hipMemset(buf, 0, slots * sizeof(T)) == hipSuccess &&
      hipEventCreate(&e0) == hipSuccess && hipEventCreate(&e1) == hipSuccess) {
    const uint32_t mask = (uint32_t)(slots - 1);
    // a warm-up launch, then the timed one
    hipLaunchKernelGGL((atomic_probe_kernel<T, OP>), dim3(blocks), dim3(threads), 0, 0, buf, mask, rounds, sink);
    bool good = hipGetLastError() == hipSuccess && hipEventRecord(e0, 0) == hipSuccess;
    if (good) hipLaunchKernelGGL((atomic_probe_kernel<T, OP>), dim3(blocks), dim3(threads), 0, 0, buf, mask, rounds, sink);
    good = good && hipGetLastError() == hipSuccess && hipEventRecord(e1, 0) == hipSuccess;
    float ms = 0;
    if (good && hipEventSynchronize(e1) == hipSuccess && hipEventElapsedTime(&ms, e0, e1) == hipSuccess && ms > 0) {
      *ops_per_s = (double)threads * blocks * rounds * HVM_PROBE_INFLIGHT / (ms * 1e-3);
      rc = HVM_HIP_OK;
    }
  }
  // teardown: a failing destroy or free has nothing left to undo
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (sink) (void)hipFree(sink);
  (void)hipFree(buf);
  return rc;
}

// The evaluator's wave primitives on n waves of inputs; per lane 7 words:
// exclusive scan, its total, max, the scan and total over LDS permutes, and
// the 64-bit sum (low, high).
__global__ __launch_bounds__(HVM_WAVE) void wave_selftest_kernel(const uint32_t* in32, const unsigned long long* in64,
                                                                  uint32_t* out) {
  const uint32_t i = blockIdx.x * HVM_WAVE + threadIdx.x;
  const uint32_t x = in32[i];
  uint32_t tot = 0, tot_ref = 0;
  const uint32_t sc = wave_scan(x, &tot);
  const uint32_t mx = wave_max(x);
  const uint32_t sc_ref = wave_scan_shfl(x, &tot_ref);
  const unsigned long long s64 = wave_sum64(in64[i]);
  uint32_t* o = out + (size_t)i * 7;
  o[0] = sc; o[1] = tot; o[2] = mx; o[3] = sc_ref; o[4] = tot_ref;
  o[5] = (uint32_t)s64; o[6] = (uint32_t)(s64 >> 32);
}

}  // namespace

extern "C" int hvm_hip_wave_selftest(uint32_t n_waves, const uint32_t* in32, const uint64_t* in64, uint32_t* out) {
  if (!in32 || !in64 || !out || n_waves == 0 || n_waves > 4096) return HVM_HIP_E_ARG;
  const size_t n = (size_t)n_waves * HVM_WAVE;
  uint32_t* d32 = nullptr;
  unsigned long long* d64 = nullptr;
  uint32_t* dout = nullptr;
  int rc = HVM_HIP_E_DEVICE;
  if (hipMalloc(&d32, n * 4) == hipSuccess && hipMalloc(&d64, n * 8) == hipSuccess &&
      hipMalloc(&dout, n * 7 * 4) == hipSuccess && hipMemcpy(d32, in32, n * 4, hipMemcpyHostToDevice) == hipSuccess &&
      hipMemcpy(d64, in64, n * 8, hipMemcpyHostToDevice) == hipSuccess) {
    hipLaunchKernelGGL(wave_selftest_kernel, dim3(n_waves), dim3(HVM_WAVE), 0, 0, d32, d64, dout);
    if (hipGetLastError() == hipSuccess && hipDeviceSynchronize() == hipSuccess &&
        hipMemcpy(out, dout, n * 7 * 4, hipMemcpyDeviceToHost) == hipSuccess)
      rc = HVM_HIP_OK;
  }
  (void)hipFree(d32);
  (void)hipFree(d64);
  (void)hipFree(dout);
  return rc;
}

extern "C" int hvm_hip_copy_rate(uint64_t bytes, double* bytes_per_s) {
  if (!bytes_per_s || bytes < (1u << 20) || bytes > (8ull << 30)) return HVM_HIP_E_ARG;
  void* src = nullptr;
  void* dst = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = HVM_HIP_E_DEVICE;
  if (hipMalloc(&src, bytes) == hipSuccess && hipMalloc(&dst, bytes) == hipSuccess &&
      hipMemset(src, 1, bytes) == hipSuccess && hipEventCreate(&e0) == hipSuccess &&
      hipEventCreate(&e1) == hipSuccess && hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice) == hipSuccess) {
    // best of 5 device-to-device copies: bytes read plus bytes written
    float best = 0;
    bool good = true;
    for (int i = 0; i < 5 && good; ++i) {
      float ms = 0;
      good = hipEventRecord(e0, 0) == hipSuccess &&
             hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, 0) == hipSuccess &&
             hipEventRecord(e1, 0) == hipSuccess && hipEventSynchronize(e1) == hipSuccess &&
             hipEventElapsedTime(&ms, e0, e1) == hipSuccess && ms > 0;
      if (good && (best == 0 || ms < best)) best = ms;
    }
    if (good) {
      *bytes_per_s = 2.0 * (double)bytes / (best * 1e-3);
      rc = HVM_HIP_OK;
    }
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  (void)hipFree(src);
  (void)hipFree(dst);
  return rc;
}

extern "C" int hvm_hip_atomic_rate(uint64_t slots, uint32_t bytes, double* ops_per_s) {
  return hvm_hip_request_rate(slots, bytes, 0, ops_per_s);
}

extern "C" int hvm_hip_request_rate(uint64_t slots, uint32_t bytes, uint32_t op, double* ops_per_s) {
  if (!ops_per_s || op > 2) return HVM_HIP_E_ARG;
  if (bytes == 4) return op == 0 ? run_probe<uint32_t, 0>(slots, ops_per_s)
                       : op == 1 ? run_probe<uint32_t, 1>(slots, ops_per_s) : run_probe<uint32_t, 2>(slots, ops_per_s);
  if (bytes == 8) return op == 0 ? run_probe<unsigned long long, 0>(slots, ops_per_s)
                       : op == 1 ? run_probe<unsigned long long, 1>(slots, ops_per_s)
                                 : run_probe<unsigned long long, 2>(slots, ops_per_s);
  return HVM_HIP_E_ARG;
}
