// LBVH construction and maintenance kernels.
//
//   scene bounds  -> 30-bit Morton codes of FP32 box centres (15 bits/axis)
//   radix sort    -> (sort.hip)
//   gather        -> boxes in Morton order + id -> sorted-position map
//   Karras 2012   -> binary hierarchy over the sorted codes ("Maximizing
//                    Parallelism in the Construction of BVHs, Octrees, and k-d
//                    Trees"), ties broken by position
//   refit         -> bottom-up union of child bounds; at every internal node the
//                    second of its two children to arrive carries on upward
//                    (atomic counter per node, device-scope acquire/release)
//
// Mutations (replacing the reference's OptiX refits, spatial_index.cuh:208-336)
// write boxes by id and in sorted position, then refit the tree.
//
// Every kernel is memory-bound integer/compare work; none uses MFMA.
#include <cfloat>

#include "morton.h"
#include "rts_internal.h"
#include "rtspatial/device/wave.h"

namespace rts {
namespace {

constexpr int kThreads = 256;

__device__ __forceinline__ bool box_dead(const float4 b) {
  // A deleted box holds NaN coordinates; any NaN makes every predicate false.
  return !(b.x == b.x) || !(b.y == b.y) || !(b.z == b.z) || !(b.w == b.w);
}

__device__ __forceinline__ float4 empty_box() {
  return make_float4(INFINITY, INFINITY, -INFINITY, -INFINITY);
}

__device__ __forceinline__ float4 unite(const float4 a, const float4 b) {
  return make_float4(fminf(a.x, b.x), fminf(a.y, b.y), fmaxf(a.z, b.z), fmaxf(a.w, b.w));
}

// ------------------------------------------------------------ bounds ----
// Bounds of the box centres, stored as (min x, min y, max x, max y).
__global__ void __launch_bounds__(kThreads) k_centre_bounds(const float4* __restrict__ boxes,
                                                            uint32_t n, float4* __restrict__ partials) {
  __shared__ float4 sh[kThreads / kWave];
  float4 acc = empty_box();
  for (uint32_t i = blockIdx.x * kThreads + threadIdx.x; i < n; i += gridDim.x * kThreads) {
    float4 b = boxes[i];
    if (box_dead(b)) continue;
    float cx = 0.5f * (b.x + b.z), cy = 0.5f * (b.y + b.w);
    acc = unite(acc, make_float4(cx, cy, cx, cy));
  }
  for (int d = 32; d >= 1; d >>= 1) {
    acc.x = fminf(acc.x, __shfl_xor(acc.x, d, 64));
    acc.y = fminf(acc.y, __shfl_xor(acc.y, d, 64));
    acc.z = fmaxf(acc.z, __shfl_xor(acc.z, d, 64));
    acc.w = fmaxf(acc.w, __shfl_xor(acc.w, d, 64));
  }
  if (lane_id() == 0) sh[threadIdx.x / kWave] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    float4 r = sh[0];
    for (int w = 1; w < kThreads / kWave; w++) r = unite(r, sh[w]);
    partials[blockIdx.x] = r;
  }
}

__global__ void __launch_bounds__(kThreads) k_reduce_bounds(const float4* __restrict__ partials,
                                                            uint32_t n, float4* __restrict__ out) {
  __shared__ float4 sh[kThreads];
  float4 acc = empty_box();
  for (uint32_t i = threadIdx.x; i < n; i += kThreads) acc = unite(acc, partials[i]);
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int s = kThreads / 2; s > 0; s >>= 1) {
    if (threadIdx.x < (uint32_t)s) sh[threadIdx.x] = unite(sh[threadIdx.x], sh[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = sh[0];
}

// ------------------------------------------------------------ morton ----
__global__ void __launch_bounds__(kThreads) k_morton_boxes(const float4* __restrict__ boxes, uint32_t n,
                                                           const float4* __restrict__ scene,
                                                           uint32_t* __restrict__ keys) {
  const float4 s = scene[0];
  for (uint32_t i = blockIdx.x * kThreads + threadIdx.x; i < n; i += gridDim.x * kThreads) {
    float4 b = boxes[i];
    keys[i] = morton_of(0.5f * (b.x + b.z), 0.5f * (b.y + b.w), s);
  }
}

__global__ void __launch_bounds__(kThreads) k_morton_points(const float2* __restrict__ pts, uint32_t n,
                                                            const float4* __restrict__ scene,
                                                            uint32_t* __restrict__ keys) {
  const float4 s = scene[0];
  for (uint32_t i = blockIdx.x * kThreads + threadIdx.x; i < n; i += gridDim.x * kThreads) {
    float2 p = pts[i];
    keys[i] = morton_of(p.x, p.y, s);
  }
}

// ------------------------------------------------------------ gather ----
__global__ void __launch_bounds__(kThreads) k_gather(const float4* __restrict__ boxes,
                                                     uint32_t* __restrict__ ids, uint32_t n,
                                                     uint32_t off, float4* __restrict__ leaf_boxes,
                                                     uint32_t* __restrict__ leaf_of_id) {
  for (uint32_t k = blockIdx.x * kThreads + threadIdx.x; k < n; k += gridDim.x * kThreads) {
    const uint32_t local = ids[k];
    if (local >= n) continue;  // unreachable for a correct sort; keeps reads in bounds
    const uint32_t id = local + off;
    if (off) ids[k] = id;
    leaf_boxes[k] = boxes[id];
    leaf_of_id[id] = off + k;
  }
}

// ------------------------------------------------------------ karras ----
// Length of the common prefix of (key, position) at sorted positions i and j;
// -1 outside [0, n).
__device__ __forceinline__ int prefix_len(const uint32_t* __restrict__ keys, int64_t n, int64_t i,
                                          int64_t j) {
  if (j < 0 || j >= n) return -1;
  uint32_t a = keys[i], b = keys[j];
  if (a == b) return 32 + __clz((uint32_t)i ^ (uint32_t)j);
  return __clz(a ^ b);
}

__global__ void __launch_bounds__(kThreads) k_karras(const uint32_t* __restrict__ keys, uint32_t n,
                                                     NodeRec* __restrict__ nodes,
                                                     uint32_t* __restrict__ parent,
                                                     uint32_t* __restrict__ flags) {
  const int64_t N = n;
  for (int64_t i = blockIdx.x * (int64_t)kThreads + threadIdx.x; i < N - 1;
       i += (int64_t)gridDim.x * kThreads) {
    const int d = (prefix_len(keys, N, i, i + 1) - prefix_len(keys, N, i, i - 1)) >= 0 ? 1 : -1;
    const int dmin = prefix_len(keys, N, i, i - d);
    int64_t lmax = 2;
    while (prefix_len(keys, N, i, i + lmax * d) > dmin) lmax <<= 1;
    int64_t l = 0;
    for (int64_t t = lmax >> 1; t >= 1; t >>= 1)
      if (prefix_len(keys, N, i, i + (l + t) * d) > dmin) l += t;
    const int64_t j = i + l * d;
    const int dnode = prefix_len(keys, N, i, j);
    int64_t s = 0, t = l;
    do {
      t = (t + 1) >> 1;
      if (prefix_len(keys, N, i, i + (s + t) * d) > dnode) s += t;
    } while (t > 1);
    const int64_t gamma = i + s * d + (d < 0 ? -1 : 0);
    const int64_t first = i < j ? i : j, last = i < j ? j : i;
    const uint32_t left = (first == gamma) ? ((uint32_t)gamma | kLeafFlag) : (uint32_t)gamma;
    const uint32_t right =
        (last == gamma + 1) ? ((uint32_t)(gamma + 1) | kLeafFlag) : (uint32_t)(gamma + 1);
    nodes[i].meta = make_uint4(left, right, (uint32_t)first, (uint32_t)last);
    const uint32_t lslot = (left & kLeafFlag) ? (uint32_t)(N - 1) + (left & ~kLeafFlag) : left;
    const uint32_t rslot = (right & kLeafFlag) ? (uint32_t)(N - 1) + (right & ~kLeafFlag) : right;
    parent[lslot] = ((uint32_t)i << 1) | 0u;
    parent[rslot] = ((uint32_t)i << 1) | 1u;
    flags[i] = 0;
  }
}

// ------------------------------------------------------------- refit ----
// Every walk ends at the root within the tree's depth (at most 65 levels for
// 30-bit keys with position tie-breaks); the level bound only guards against
// a corrupt parent array, and reports it through `error`.
constexpr int kMaxLevels = 128;
// Leaves per workgroup (chunk). The single workgroup of k_refit_top walks
// every node that spans chunks, so larger chunks shorten it, while each
// thread's walks run one after another and fewer workgroups share the GPU.
// Measured on 10M boxes (tree phase of a build, refit included): 2048 leaves
// 3.20 ms, 4096 2.70, 8192 2.36, 12288 2.48, 16384 2.64, 32768 3.60; a tree
// of 10k-100k boxes (a delta tree) wants 2048-leaf chunks (0.25 ms per
// delta build instead of 0.55 ms with 8192).
__host__ __device__ constexpr uint32_t refit_chunk(uint32_t n) {
  return n >= (1u << 22) ? 8192u : 2048u;
}

// Workgroup b walks up from leaves [b*C, (b+1)*C), C = refit_chunk(n). A node whose leaf
// range lies inside that chunk has both children finished by this workgroup,
// so its arrival counter only needs workgroup-scope ordering (no L2
// write-back or invalidation); only nodes that straddle chunks go to
// k_refit_top. Each child slot also receives the
// number of live (not deleted) leaves below it.
__global__ void __launch_bounds__(kThreads) k_refit(const float4* __restrict__ leaf_boxes, uint32_t n,
                                                    const uint32_t* __restrict__ parent,
                                                    NodeRec* nodes, float4* __restrict__ root_box,
                                                    uint32_t* __restrict__ root_live,
                                                    uint32_t* flags, uint32_t* error,
                                                    uint32_t* boundary) {
  const uint32_t chunk = refit_chunk(n);
  const uint32_t c0 = blockIdx.x * chunk;
  const uint32_t c1 = (c0 + chunk < n ? c0 + chunk : n) - 1;  // inclusive
  for (uint32_t r = 0; r < chunk / kThreads; r++) {
    const uint32_t k = c0 + r * kThreads + threadIdx.x;
    if (k >= n) break;
    float4 b = leaf_boxes[k];
    uint32_t live = 1;
    if (box_dead(b)) {
      b = empty_box();
      live = 0;
    } else {
      // Bounds cover the box's extent even if its corners are swapped, so a
      // node inside a query implies every live box below it matches.
      b = make_float4(fminf(b.x, b.z), fminf(b.y, b.w), fmaxf(b.x, b.z), fmaxf(b.y, b.w));
    }
    if (n == 1) {
      root_box[0] = b;
      root_live[0] = live;
      continue;
    }
    uint32_t code = parent[(n - 1) + k];
    for (int level = 0;; level++) {
      const uint32_t p = code >> 1, side = code & 1u;
      if (level >= kMaxLevels || p >= n - 1) {
        atomicOr(error, 2u);
        break;
      }
      NodeRec* np = nodes + p;
      np->box[side] = b;
      if (side) np->live.y = live; else np->live.x = live;
      const uint4 mp = np->meta;
      if (!(mp.z >= c0 && mp.w <= c1)) {
        // The node spans chunks: k_refit_top finishes it (and everything
        // above) after this kernel, so no agent-scope ordering is needed.
        const uint32_t slot = atomicAdd(&boundary[0], 1u);
        if (1 + (uint64_t)slot < refit_boundary_words(n)) boundary[1 + slot] = code;
        else atomicOr(error, 2u);
        break;
      }
      // Release our child's bounds; acquire the sibling's if we arrive second.
      const uint32_t old =
          __hip_atomic_fetch_add(&flags[p], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP);
      if (old == 0) break;
      const float4 o = np->box[side ^ 1u];
      live += side ? np->live.x : np->live.y;
      b = unite(b, o);
      if (p == 0) {
        root_box[0] = b;
        root_live[0] = live;
        break;
      }
      code = parent[p];
    }
  }
}

// The nodes that span chunks, bottom-up, in one workgroup: each boundary
// code names a child slot that k_refit filled; the second arrival at a node
// unites both slots and fills the node's slot in its parent.
constexpr int kRefitTopThreads = 1024;

__global__ void __launch_bounds__(kRefitTopThreads) k_refit_top(const uint32_t* __restrict__ parent,
                                                              NodeRec* nodes, uint32_t n,
                                                              float4* __restrict__ root_box,
                                                              uint32_t* __restrict__ root_live,
                                                              uint32_t* flags, uint32_t* error,
                                                              const uint32_t* boundary) {
  const uint64_t cap = refit_boundary_words(n) - 1;
  const uint32_t cnt = boundary[0] < cap ? boundary[0] : (uint32_t)cap;
  for (uint32_t j = threadIdx.x; j < cnt; j += kRefitTopThreads) {
    uint32_t code = boundary[1 + j];
    for (int level = 0;; level++) {
      const uint32_t p = code >> 1;
      if (level >= kMaxLevels || p >= n - 1) {
        atomicOr(error, 2u);
        break;
      }
      NodeRec* np = nodes + p;
      const uint32_t old =
          __hip_atomic_fetch_add(&flags[p], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP);
      if (old == 0) break;
      const float4 b = unite(np->box[0], np->box[1]);
      const uint32_t live = np->live.x + np->live.y;
      if (p == 0) {
        root_box[0] = b;
        root_live[0] = live;
        break;
      }
      code = parent[p];
      NodeRec* pp = nodes + (code >> 1);
      pp->box[code & 1u] = b;
      if (code & 1u) pp->live.y = live; else pp->live.x = live;
    }
  }
}

__global__ void __launch_bounds__(kThreads) k_clear_flags(uint32_t* flags, uint32_t m) {
  for (uint32_t i = blockIdx.x * kThreads + threadIdx.x; i < m; i += gridDim.x * kThreads) flags[i] = 0;
}

// ------------------------------------------------------------ cost ------
// Sum of half-perimeters of all nodes below the root, divided by the root's:
// the expected number of node visits of small random range queries, up to a
// constant. Non-finite (empty) boxes contribute nothing.
__global__ void __launch_bounds__(kThreads) k_tree_cost(const NodeRec* __restrict__ nodes, uint32_t m,
                                                        double* __restrict__ out) {
  __shared__ double sh[kThreads / kWave];
  double acc = 0.0;
  for (uint32_t i = blockIdx.x * kThreads + threadIdx.x; i < m; i += gridDim.x * kThreads) {
    float4 b = nodes[i >> 1].box[i & 1];
    double hp = (double)(b.z - b.x) + (double)(b.w - b.y);
    if (hp >= 0.0 && hp < 1e300) acc += hp;
  }
  acc = wave_sum_f64(acc);
  if (lane_id() == 0) sh[threadIdx.x / kWave] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0;
    for (int w = 0; w < kThreads / kWave; w++) t += sh[w];
    atomicAdd(out, t);
  }
}

__global__ void k_cost_finish(const float4* __restrict__ root_box, double* out) {
  float4 r = root_box[0];
  double hp = (double)(r.z - r.x) + (double)(r.w - r.y);
  out[1] = (hp > 0.0 && hp < 1e300) ? (out[0] + hp) / hp : 1.0;
}

// --------------------------------------------------------- mutations ----
__global__ void __launch_bounds__(kThreads) k_delete(const uint64_t* __restrict__ ids, uint64_t m,
                                                     float4* __restrict__ boxes, uint64_t n_boxes,
                                                     float4* __restrict__ leaf_boxes,
                                                     const uint32_t* __restrict__ leaf_of_id,
                                                     uint64_t n_tree, Box64* __restrict__ boxes64) {
  const float nan = __builtin_nanf("");
  const float4 dead = make_float4(nan, nan, nan, nan);
  const double dnan = __builtin_nan("");
  for (uint64_t i = blockIdx.x * (uint64_t)kThreads + threadIdx.x; i < m;
       i += (uint64_t)gridDim.x * kThreads) {
    uint64_t id = ids[i];
    if (id >= n_boxes) continue;  // out-of-range ids are ignored
    boxes[id] = dead;
    if (boxes64) boxes64[id] = Box64{dnan, dnan, dnan, dnan};
    if (id < n_tree) leaf_boxes[leaf_of_id[id]] = dead;
  }
}

__global__ void __launch_bounds__(kThreads) k_update(const rts_update* __restrict__ ups, uint64_t m,
                                                     float4* __restrict__ boxes, uint64_t n_boxes,
                                                     float4* __restrict__ leaf_boxes,
                                                     const uint32_t* __restrict__ leaf_of_id,
                                                     uint64_t n_tree) {
  for (uint64_t i = blockIdx.x * (uint64_t)kThreads + threadIdx.x; i < m;
       i += (uint64_t)gridDim.x * kThreads) {
    const rts_update u = ups[i];
    if (u.id >= n_boxes) continue;
    const float4 b = make_float4(u.box.x0, u.box.y0, u.box.x1, u.box.y1);
    boxes[u.id] = b;
    if (u.id < n_tree) leaf_boxes[leaf_of_id[u.id]] = b;
  }
}

uint32_t grid_for(uint64_t n, uint32_t cap = 8192) {
  uint64_t g = (n + kThreads - 1) / kThreads;
  if (g < 1) g = 1;
  return (uint32_t)(g > cap ? cap : g);
}

}  // namespace

uint32_t scene_bounds_partials(uint32_t n) { return grid_for(n, 1024); }

hipError_t launch_scene_bounds(const float4* boxes, uint32_t n, float4* partials,
                               uint32_t n_partials, float4* scene, hipStream_t stream) {
  hipLaunchKernelGGL(k_centre_bounds, dim3(n_partials), dim3(kThreads), 0, stream, boxes, n,
                     partials);
  hipLaunchKernelGGL(k_reduce_bounds, dim3(1), dim3(kThreads), 0, stream, partials, n_partials,
                     scene);
  return hipGetLastError();
}

hipError_t launch_morton_boxes(const float4* boxes, uint32_t n, const float4* scene, uint32_t* keys,
                               hipStream_t stream) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_morton_boxes, dim3(grid_for(n)), dim3(kThreads), 0, stream, boxes, n, scene,
                     keys);
  return hipGetLastError();
}

hipError_t launch_morton_points(const float2* pts, uint32_t n, const float4* scene, uint32_t* keys,
                                hipStream_t stream) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_morton_points, dim3(grid_for(n)), dim3(kThreads), 0, stream, pts, n, scene,
                     keys);
  return hipGetLastError();
}

hipError_t launch_gather_leaves(const float4* boxes, uint32_t* sorted_ids, uint32_t n, uint32_t off,
                                float4* leaf_boxes, uint32_t* leaf_of_id, hipStream_t stream) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_gather, dim3(grid_for(n)), dim3(kThreads), 0, stream, boxes, sorted_ids, n,
                     off, leaf_boxes, leaf_of_id);
  return hipGetLastError();
}

// Every stride-th element of `src` into dst[0, m): the autotuner's query
// sample. Elements are 8-byte float points, 16-byte float envelopes or double
// points, or 32-byte double envelopes (two 16-byte words).
template <typename T, int W>
__global__ void __launch_bounds__(kThreads) k_gather_strided(const T* __restrict__ src, uint64_t stride,
                                                             T* __restrict__ dst, uint32_t m) {
  for (uint32_t i = blockIdx.x * kThreads + threadIdx.x; i < m; i += gridDim.x * kThreads)
#pragma unroll
    for (int w = 0; w < W; w++) dst[(uint64_t)i * W + w] = src[(uint64_t)i * stride * W + w];
}

hipError_t launch_gather_strided(const void* src, uint32_t elem_bytes, uint64_t stride, void* dst,
                                 uint32_t m, hipStream_t stream) {
  if (m == 0) return hipSuccess;
  if (elem_bytes == 8)
    hipLaunchKernelGGL((k_gather_strided<float2, 1>), dim3(grid_for(m)), dim3(kThreads), 0, stream,
                       static_cast<const float2*>(src), stride, static_cast<float2*>(dst), m);
  else if (elem_bytes == 16)
    hipLaunchKernelGGL((k_gather_strided<float4, 1>), dim3(grid_for(m)), dim3(kThreads), 0, stream,
                       static_cast<const float4*>(src), stride, static_cast<float4*>(dst), m);
  else if (elem_bytes == 32)
    hipLaunchKernelGGL((k_gather_strided<float4, 2>), dim3(grid_for(m)), dim3(kThreads), 0, stream,
                       static_cast<const float4*>(src), stride, static_cast<float4*>(dst), m);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

hipError_t launch_karras(const uint32_t* sorted_keys, uint32_t n, NodeRec* nodes, uint32_t* parent,
                         uint32_t* flags, hipStream_t stream) {
  if (n < 2) return hipSuccess;
  hipLaunchKernelGGL(k_karras, dim3(grid_for(n - 1)), dim3(kThreads), 0, stream, sorted_keys, n,
                     nodes, parent, flags);
  return hipGetLastError();
}

hipError_t launch_refit(const RefitArgs& a, hipStream_t stream) {
  const uint32_t n = a.n;
  if (n == 0) return hipSuccess;
  if (n > 1)
    hipLaunchKernelGGL(k_clear_flags, dim3(grid_for(n - 1)), dim3(kThreads), 0, stream, a.flags,
                       n - 1);
  hipError_t e = launch_fill_u32(a.boundary, 0u, 1, stream);
  if (e != hipSuccess) return e;
  const uint32_t blocks = (n + refit_chunk(n) - 1) / refit_chunk(n);
  hipLaunchKernelGGL(k_refit, dim3(blocks), dim3(kThreads), 0, stream, a.leaf_boxes, n, a.parent,
                     a.nodes, a.root_box, a.root_live, a.flags, a.error, a.boundary);
  if (n > 1)
    hipLaunchKernelGGL(k_refit_top, dim3(1), dim3(kRefitTopThreads), 0, stream, a.parent, a.nodes,
                       n, a.root_box, a.root_live, a.flags, a.error, a.boundary);
  return hipGetLastError();
}

hipError_t launch_tree_cost(const NodeRec* nodes, const float4* root_box, uint32_t n,
                            double* out, hipStream_t stream) {
  hipError_t e = launch_fill_u32(reinterpret_cast<uint32_t*>(out), 0u, 4, stream);
  if (e != hipSuccess) return e;
  if (n > 1)
    hipLaunchKernelGGL(k_tree_cost, dim3(grid_for(2 * (uint64_t)(n - 1), 2048)), dim3(kThreads), 0,
                       stream, nodes, 2 * (n - 1), out);
  hipLaunchKernelGGL(k_cost_finish, dim3(1), dim3(1), 0, stream, root_box, out);
  return hipGetLastError();
}

hipError_t launch_delete(const uint64_t* ids, uint64_t n_ids, float4* boxes, uint64_t n_boxes,
                         float4* leaf_boxes, const uint32_t* leaf_of_id, uint64_t n_tree,
                         Box64* boxes64, hipStream_t stream) {
  if (n_ids == 0) return hipSuccess;
  hipLaunchKernelGGL(k_delete, dim3(grid_for(n_ids)), dim3(kThreads), 0, stream, ids, n_ids, boxes,
                     n_boxes, leaf_boxes, leaf_of_id, n_tree, boxes64);
  return hipGetLastError();
}

hipError_t launch_update(const rts_update* ups, uint64_t n_ups, float4* boxes, uint64_t n_boxes,
                         float4* leaf_boxes, const uint32_t* leaf_of_id, uint64_t n_tree,
                         hipStream_t stream) {
  if (n_ups == 0) return hipSuccess;
  hipLaunchKernelGGL(k_update, dim3(grid_for(n_ups)), dim3(kThreads), 0, stream, ups, n_ups, boxes,
                     n_boxes, leaf_boxes, leaf_of_id, n_tree);
  return hipGetLastError();
}

}  // namespace rts
