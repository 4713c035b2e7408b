// Double-coordinate support (SpatialIndex<double, 2>): conversion of double
// boxes to the float boxes the tree is built from, double updates, and Morton
// keys of double queries. The tree, its build and refit, and every traversal
// step above the leaves stay in FP32; see round.h for the rounding rules.
//
// All kernels are streaming passes (32-byte double box in, 16-byte float box
// out), HBM-bound.
#include "morton.h"
#include "rtspatial/device/round.h"
#include "rts_internal.h"

namespace rts {
namespace {

constexpr int kThreads = 256;

uint32_t grid_for(uint64_t n, uint32_t cap = 8192) {
  uint64_t g = (n + kThreads - 1) / kThreads;
  if (g < 1) g = 1;
  return (uint32_t)(g > cap ? cap : g);
}

__device__ __forceinline__ float4 round_box(const Box64& b, int rounding) {
  if (rounding == kRoundNearest)
    return make_float4((float)b.x0, (float)b.y0, (float)b.x1, (float)b.y1);
  return make_float4(f32_rd(b.x0), f32_rd(b.y0), f32_ru(b.x1), f32_ru(b.y1));
}

__global__ void __launch_bounds__(kThreads) k_round_boxes64(const Box64* __restrict__ src, uint64_t n,
                                                            float4* __restrict__ dst, int rounding) {
  for (uint64_t i = blockIdx.x * (uint64_t)kThreads + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * kThreads)
    dst[i] = round_box(src[i], rounding);
}

__global__ void __launch_bounds__(kThreads) k_round_points64(const double2* __restrict__ src,
                                                             uint64_t n, float2* __restrict__ dst) {
  for (uint64_t i = blockIdx.x * (uint64_t)kThreads + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * kThreads) {
    const double2 p = src[i];
    dst[i] = make_float2((float)p.x, (float)p.y);
  }
}

__global__ void __launch_bounds__(kThreads) k_update64(const rts_update_d* __restrict__ ups, uint64_t m,
                                                       Box64* __restrict__ boxes64,
                                                       float4* __restrict__ boxes, uint64_t n_boxes,
                                                       float4* __restrict__ leaf_boxes,
                                                       const uint32_t* __restrict__ leaf_of_id,
                                                       uint64_t n_tree, int rounding) {
  for (uint64_t i = blockIdx.x * (uint64_t)kThreads + threadIdx.x; i < m;
       i += (uint64_t)gridDim.x * kThreads) {
    const rts_update_d u = ups[i];
    if (u.id >= n_boxes) continue;  // out-of-range ids are ignored
    const Box64 d{u.box.x0, u.box.y0, u.box.x1, u.box.y1};
    const float4 f = round_box(d, rounding);
    boxes64[u.id] = d;
    boxes[u.id] = f;
    if (u.id < n_tree) leaf_boxes[leaf_of_id[u.id]] = f;
  }
}

// Query ordering only needs nearby queries to get nearby keys: the centre is
// taken in double and rounded to nearest.
__global__ void __launch_bounds__(kThreads) k_morton64(const void* __restrict__ queries, int points,
                                                       uint32_t n, const float4* __restrict__ scene,
                                                       uint32_t* __restrict__ keys) {
  const float4 s = scene[0];
  for (uint32_t i = blockIdx.x * kThreads + threadIdx.x; i < n; i += gridDim.x * kThreads) {
    double cx, cy;
    if (points) {
      const double2 p = static_cast<const double2*>(queries)[i];
      cx = p.x;
      cy = p.y;
    } else {
      const Box64 b = static_cast<const Box64*>(queries)[i];
      cx = 0.5 * (b.x0 + b.x1);
      cy = 0.5 * (b.y0 + b.y1);
    }
    keys[i] = morton_of((float)cx, (float)cy, s);
  }
}

}  // namespace

hipError_t launch_round_boxes64(const Box64* src, uint64_t n, float4* dst, int rounding,
                                hipStream_t stream) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_round_boxes64, dim3(grid_for(n)), dim3(kThreads), 0, stream, src, n, dst,
                     rounding);
  return hipGetLastError();
}

hipError_t launch_round_queries64(const void* src, bool points, uint64_t n, void* dst,
                                  hipStream_t stream) {
  if (n == 0) return hipSuccess;
  if (points)
    hipLaunchKernelGGL(k_round_points64, dim3(grid_for(n)), dim3(kThreads), 0, stream,
                       static_cast<const double2*>(src), n, static_cast<float2*>(dst));
  else
    hipLaunchKernelGGL(k_round_boxes64, dim3(grid_for(n)), dim3(kThreads), 0, stream,
                       static_cast<const Box64*>(src), n, static_cast<float4*>(dst),
                       (int)kRoundNearest);
  return hipGetLastError();
}

hipError_t launch_update64(const rts_update_d* ups, uint64_t n_ups, Box64* boxes64, float4* boxes,
                           uint64_t n_boxes, float4* leaf_boxes, const uint32_t* leaf_of_id,
                           uint64_t n_tree, int rounding, hipStream_t stream) {
  if (n_ups == 0) return hipSuccess;
  hipLaunchKernelGGL(k_update64, dim3(grid_for(n_ups)), dim3(kThreads), 0, stream, ups, n_ups,
                     boxes64, boxes, n_boxes, leaf_boxes, leaf_of_id, n_tree, rounding);
  return hipGetLastError();
}

hipError_t launch_morton64(const void* queries, bool points, uint32_t n, const float4* scene,
                           uint32_t* keys, hipStream_t stream) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_morton64, dim3(grid_for(n)), dim3(kThreads), 0, stream, queries,
                     points ? 1 : 0, n, scene, keys);
  return hipGetLastError();
}

}  // namespace rts
