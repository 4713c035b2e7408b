// Wavefront (wave64) helpers for CDNA.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace rts {

__device__ __forceinline__ uint32_t lane_id() {
  return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
}

// Number of set bits of `mask` in lanes below the calling lane.
__device__ __forceinline__ uint32_t mask_rank(uint64_t mask) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                   __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }

__device__ __forceinline__ uint32_t readlane_u32(uint32_t v, uint32_t lane) {
  return (uint32_t)__builtin_amdgcn_readlane((int)v, (int)lane);
}

__device__ __forceinline__ uint64_t readlane_u64(uint64_t v, uint32_t lane) {
  uint32_t lo = readlane_u32((uint32_t)v, lane);
  uint32_t hi = readlane_u32((uint32_t)(v >> 32), lane);
  return ((uint64_t)hi << 32) | lo;
}

// Inclusive prefix sum over the 64 lanes (all lanes must be active).
__device__ __forceinline__ uint32_t wave_inclusive_sum(uint32_t v) {
  const uint32_t lane = lane_id();
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    uint32_t o = __shfl_up(v, d, 64);
    if (lane >= (uint32_t)d) v += o;
  }
  return v;
}

// Inclusive prefix maximum over the 64 lanes (all lanes must be active).
__device__ __forceinline__ uint32_t wave_inclusive_max(uint32_t v) {
  const uint32_t lane = lane_id();
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    uint32_t o = __shfl_up(v, d, 64);
    if (lane >= (uint32_t)d) v = v > o ? v : o;
  }
  return v;
}

__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t v) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
  return v;
}

__device__ __forceinline__ double wave_sum_f64(double v) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
  return v;
}

// Device-scope relaxed atomics for flags that carry their own payload.
__device__ __forceinline__ uint32_t load_relaxed_agent(const uint32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void store_relaxed_agent(uint32_t* p, uint32_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

}  // namespace rts
