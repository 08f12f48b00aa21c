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

// Cross-lane arithmetic on DPP (data-parallel primitives) moves, which run in
// the vector ALU: row_shr:n shifts within each 16-lane row (lanes shifted in
// from outside the row read 0), and readlane takes the four row totals into
// scalar registers.  The same operations over LDS permutes (__shfl) cost six
// dependent LDS round trips each; the kernels run several per step.  Every
// lane of the wave must be active.
template <int CTRL>
__device__ __forceinline__ uint32_t dpp_shr(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, 0xf, 0xf, true);
}
#define HVM_ROW_SHR(n) (0x110 + (n))

__device__ __forceinline__ uint32_t lane_u32(uint32_t v, int lane) {
  return (uint32_t)__builtin_amdgcn_readlane((int)v, lane);
}

// Inclusive prefix sum within each 16-lane row.
__device__ __forceinline__ uint32_t row_scan(uint32_t v) {
  v += dpp_shr<HVM_ROW_SHR(1)>(v);
  v += dpp_shr<HVM_ROW_SHR(2)>(v);
  v += dpp_shr<HVM_ROW_SHR(4)>(v);
  v += dpp_shr<HVM_ROW_SHR(8)>(v);
  return v;
}

// Exclusive prefix sum over the wave; *total receives the wave sum.
__device__ __forceinline__ uint32_t wave_scan(uint32_t x, uint32_t* total) {
  const uint32_t v = row_scan(x);
  const uint32_t r0 = lane_u32(v, 15), r1 = lane_u32(v, 31), r2 = lane_u32(v, 47), r3 = lane_u32(v, 63);
  const uint32_t row = lane_id() >> 4;
  const uint32_t before = (row >= 1 ? r0 : 0u) + (row >= 2 ? r1 : 0u) + (row >= 3 ? r2 : 0u);
  *total = r0 + r1 + r2 + r3;
  return v + before - x;
}

__device__ __forceinline__ uint32_t wave_max(uint32_t x) {
  uint32_t v = x;
  v = max(v, dpp_shr<HVM_ROW_SHR(1)>(v));
  v = max(v, dpp_shr<HVM_ROW_SHR(2)>(v));
  v = max(v, dpp_shr<HVM_ROW_SHR(4)>(v));
  v = max(v, dpp_shr<HVM_ROW_SHR(8)>(v));
  return max(max(lane_u32(v, 15), lane_u32(v, 31)), max(lane_u32(v, 47), lane_u32(v, 63)));
}

// Sum over the wave of a 64-bit value (every lane receives it): the row
// prefix sums of both halves, with the carries of the low half.
__device__ __forceinline__ unsigned long long wave_sum64(unsigned long long x) {
  uint32_t lo = (uint32_t)x, hi = (uint32_t)(x >> 32);
#define HVM_ADD64_SHR(n)                          \
  {                                               \
    const uint32_t l2 = dpp_shr<HVM_ROW_SHR(n)>(lo); \
    const uint32_t h2 = dpp_shr<HVM_ROW_SHR(n)>(hi); \
    const uint32_t nl = lo + l2;                  \
    hi = hi + h2 + (nl < lo ? 1u : 0u);           \
    lo = nl;                                      \
  }
  HVM_ADD64_SHR(1)
  HVM_ADD64_SHR(2)
  HVM_ADD64_SHR(4)
  HVM_ADD64_SHR(8)
#undef HVM_ADD64_SHR
  unsigned long long t = 0;
  for (int r = 15; r < 64; r += 16)
    t += ((unsigned long long)lane_u32(hi, r) << 32) | (unsigned long long)lane_u32(lo, r);
  return t;
}

// The same operations over LDS permutes, for the self-test of the above.
__device__ __forceinline__ uint32_t wave_scan_shfl(uint32_t x, uint32_t* total) {
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
