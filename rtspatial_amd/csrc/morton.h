// 30-bit Morton codes (15 bits per axis) of points inside the scene bounds,
// shared by the tree build (box centres) and the query ordering.
#pragma once

#include "rts_internal.h"

namespace rts {

__device__ __forceinline__ uint32_t spread15(uint32_t v) {
  v &= 0x7FFF;
  v = (v | (v << 8)) & 0x00FF00FFu;
  v = (v | (v << 4)) & 0x0F0F0F0Fu;
  v = (v | (v << 2)) & 0x33333333u;
  v = (v | (v << 1)) & 0x55555555u;
  return v;
}

__device__ __forceinline__ uint32_t quantize(float c, float lo, float scale) {
  float t = (c - lo) * scale;
  t = fminf(fmaxf(t, 0.0f), 32767.0f);
  return (uint32_t)t;
}

__device__ __forceinline__ uint32_t morton_of(float cx, float cy, const float4 scene) {
  if (!(cx == cx) || !(cy == cy)) return (1u << kMortonBits) - 1;  // deleted: sort last
  float sx = scene.z - scene.x, sy = scene.w - scene.y;
  float kx = sx > 0.0f ? 32767.0f / sx : 0.0f;
  float ky = sy > 0.0f ? 32767.0f / sy : 0.0f;
  return (spread15(quantize(cy, scene.y, ky)) << 1) | spread15(quantize(cx, scene.x, kx));
}

}  // namespace rts
