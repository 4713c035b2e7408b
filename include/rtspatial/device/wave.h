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

// DPP lane moves (GFX9 data-parallel primitives; VALU operations, no LDS).
// Lanes whose source lies outside the row (row_shr) or whose row is masked
// off (row_bcast) read 0.
enum : int {
  kDppRowShr1 = 0x111,
  kDppRowShr2 = 0x112,
  kDppRowShr4 = 0x114,
  kDppRowShr8 = 0x118,
  kDppWaveShr1 = 0x138,
  kDppRowBcast15 = 0x142,  // lane 15 of each row to the next row
  kDppRowBcast31 = 0x143,  // lane 31 to rows 2 and 3
};
template <int CTRL, int ROW_MASK = 0xf, bool BOUND_CTRL = true>
__device__ __forceinline__ uint32_t dpp_move(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, ROW_MASK, 0xf, BOUND_CTRL);
}

// Inclusive prefix sum over the 64 lanes (all lanes must be active): shifts
// by 1, 2, 4, 8 within each row of 16, then row 0's and row 2's totals into
// rows 1 and 3, then lane 31's into rows 2 and 3.
__device__ __forceinline__ uint32_t wave_inclusive_sum(uint32_t v) {
  v += dpp_move<kDppRowShr1>(v);
  v += dpp_move<kDppRowShr2>(v);
  v += dpp_move<kDppRowShr4>(v);
  v += dpp_move<kDppRowShr8>(v);
  v += dpp_move<kDppRowBcast15, 0xa, false>(v);
  v += dpp_move<kDppRowBcast31, 0xc, false>(v);
  return v;
}

// Inclusive prefix maximum over the 64 lanes (all lanes must be active),
// the same pattern as wave_inclusive_sum.
__device__ __forceinline__ uint32_t wave_inclusive_max(uint32_t v) {
  uint32_t o;
  o = dpp_move<kDppRowShr1>(v);
  v = v > o ? v : o;
  o = dpp_move<kDppRowShr2>(v);
  v = v > o ? v : o;
  o = dpp_move<kDppRowShr4>(v);
  v = v > o ? v : o;
  o = dpp_move<kDppRowShr8>(v);
  v = v > o ? v : o;
  o = dpp_move<kDppRowBcast15, 0xa, false>(v);
  v = v > o ? v : o;
  o = dpp_move<kDppRowBcast31, 0xc, false>(v);
  return v > o ? v : o;
}

// The value of the lane below (lane 0 reads 0); all lanes must be active.
__device__ __forceinline__ uint32_t wave_shift_up1(uint32_t v) {
  return dpp_move<kDppWaveShr1>(v);
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
