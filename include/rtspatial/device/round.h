// Directed rounding of doubles to floats, and the float neighbours of a float.
//
// The tree of a double index holds float bounds rounded outward, so that the
// float box of a double box d contains it: (x0, y0) rounded down, (x1, y1)
// rounded up (the reference widens by two ulps for its OptiX AABBs,
// geom/envelope.cuh:131-143 with utils/helpers.h:26-41; one directed rounding
// is the tightest bound). For a float f and a double q:
//   q >= f  <=>  down(q) >= f        q <= f  <=>  up(q) <= f
// so a float bound compared with a double query coordinate is decided exactly
// by the query rounded in the matching direction.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace rts {

// Next float above / below f. Infinities step to the largest finite float
// (inward); +inf up and -inf down give NaN, which fails every comparison;
// NaN stays NaN.
__host__ __device__ __forceinline__ float f32_up(float f) {
  if (!(f == f)) return f;
  if (f == 0.0f) return __builtin_bit_cast(float, 1u);  // +/-0 -> smallest subnormal
  const uint32_t u = __builtin_bit_cast(uint32_t, f);
  return __builtin_bit_cast(float, f > 0.0f ? u + 1u : u - 1u);
}
__host__ __device__ __forceinline__ float f32_down(float f) {
  if (!(f == f)) return f;
  if (f == 0.0f) return __builtin_bit_cast(float, 0x80000001u);
  const uint32_t u = __builtin_bit_cast(uint32_t, f);
  return __builtin_bit_cast(float, f > 0.0f ? u - 1u : u + 1u);
}

// The largest float <= d / the smallest float >= d (NaN stays NaN). The
// conversion rounds to nearest; one step corrects it when it went the wrong way.
__host__ __device__ __forceinline__ float f32_rd(double d) {
  const float f = (float)d;
  return (double)f > d ? f32_down(f) : f;
}
__host__ __device__ __forceinline__ float f32_ru(double d) {
  const float f = (float)d;
  return (double)f < d ? f32_up(f) : f;
}

}  // namespace rts
