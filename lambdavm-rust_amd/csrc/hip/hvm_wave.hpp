// hvm_wave.hpp — wave-level primitives shared by the evaluator kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "hvm_dev.hpp"

// ---------------------------------------------------------------------------
// Wave primitives
// ---------------------------------------------------------------------------

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

__device__ __forceinline__ uint32_t rank_in(unsigned long long mask) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

// Exclusive prefix sum over the wave; *total receives the wave sum.
__device__ __forceinline__ uint32_t wave_scan(uint32_t x, uint32_t* total) {
  uint32_t v = x;
  uint32_t lane = lane_id();
#pragma unroll
  for (uint32_t d = 1; d < HVM_WAVE; d <<= 1) {
    uint32_t t = __shfl_up(v, d, HVM_WAVE);
    if (lane >= d) v += t;
  }
  *total = __shfl(v, HVM_WAVE - 1, HVM_WAVE);
  return v - x;
}

__device__ __forceinline__ uint32_t wave_max(uint32_t x) {
#pragma unroll
  for (uint32_t d = 1; d < HVM_WAVE; d <<= 1) {
    uint32_t t = __shfl_xor(x, d, HVM_WAVE);
    x = x > t ? x : t;
  }
  return x;
}

__device__ __forceinline__ unsigned long long wave_sum64(unsigned long long x) {
#pragma unroll
  for (uint32_t d = 1; d < HVM_WAVE; d <<= 1) x += __shfl_xor(x, d, HVM_WAVE);
  return x;
}

