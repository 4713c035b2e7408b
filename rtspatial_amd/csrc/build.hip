// LBVH construction and maintenance kernels.
//
//   scene bounds  -> 30-bit Morton codes of FP32 box centres (15 bits/axis)
//   radix sort    -> (sort.hip)
//   Karras 2012   -> binary hierarchy over the sorted codes ("Maximizing
//                    Parallelism in the Construction of BVHs, Octrees, and k-d
//                    Trees"), ties broken by position; small nodes answered
//                    from an LDS window of the keys
//   refit         -> bottom-up union of child bounds in three levels (LDS
//                    node state per chunk, then the nodes that span chunks);
//                    a build's refit also gathers the boxes into Morton order
//                    and the id -> sorted-position map
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

// ------------------------------------------------------------ karras ----
// Record meta of internal node i with far end j and split gamma: {left ref,
// right ref, first leaf, last leaf}; a child is a leaf when its range is the
// one position gamma (left) or gamma + 1 (right).
__device__ __forceinline__ uint4 node_meta(uint32_t i, uint32_t j, uint32_t gamma) {
  const uint32_t first = i < j ? i : j, last = i < j ? j : i;
  return make_uint4(first == gamma ? (gamma | kLeafFlag) : gamma,
                    last == gamma + 1 ? ((gamma + 1) | kLeafFlag) : gamma + 1, first, last);
}

// Length of the common prefix of (key, position) at sorted positions i and j;
// -1 outside [0, n).
__device__ __forceinline__ int prefix_len(const uint32_t* __restrict__ keys, int64_t n, int64_t i,
                                          int64_t j) {
  if (j < 0 || j >= n) return -1;
  uint32_t a = keys[i], b = keys[j];
  if (a == b) return 32 + __clz((uint32_t)i ^ (uint32_t)j);
  return __clz(a ^ b);
}

// Karras 2012 over the sorted (key, position) pairs: node i's direction, its
// range's far end j by an exponential then a binary search, and its split
// gamma by a binary search. The refit writes each node's record whole (one
// 64-byte line from one thread: partial-line writes from two kernels cost
// more than the line): this kernel stores j in the compact node_end array,
// from which (with the arriving child) the refit rebuilds the record's meta,
// and writes the children's parent codes. Only a node whose range spans
// refit chunks -- finished by refit levels 2 and 3, which read meta from the
// record -- gets its meta here.
__global__ void __launch_bounds__(kThreads) k_karras(const uint32_t* __restrict__ keys, uint32_t n,
                                                     NodeRec* __restrict__ nodes,
                                                     uint32_t* __restrict__ parent,
                                                     uint32_t* __restrict__ node_end,
                                                     uint32_t chunk_log2) {
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
    const uint4 m = node_meta((uint32_t)i, (uint32_t)j, (uint32_t)gamma);
    node_end[i] = (uint32_t)j;
    const uint32_t lslot = (m.x & kLeafFlag) ? (uint32_t)(N - 1) + (m.x & ~kLeafFlag) : m.x;
    const uint32_t rslot = (m.y & kLeafFlag) ? (uint32_t)(N - 1) + (m.y & ~kLeafFlag) : m.y;
    // kSpanBit: node i's leaf range crosses a refit chunk boundary.
    const uint32_t span = (m.z >> chunk_log2) != (m.w >> chunk_log2) ? kSpanBit : 0u;
    if (span) {
      nodes[i].meta = m;
      nodes[i].live.z = (uint32_t)gamma;  // the split, read by the wave walks
    }
    parent[lslot] = ((uint32_t)i << 1) | 0u | span;
    parent[rslot] = ((uint32_t)i << 1) | 1u | span;
  }
}

// ------------------------------------------------------------- refit ----
// Bottom-up refit in three levels, so that no level needs ordering across
// workgroups (agent-scope acquire/release across the XCDs' non-coherent L2s
// made a one-kernel refit 2.6x slower):
//
//   1. k_refit_chunk: workgroup c takes leaves [c*C, (c+1)*C) (C = refit
//      chunk). Every node whose leaf range lies in that chunk has its index
//      in the chunk too (a Karras node is one end of its range), so its
//      arrival counter, live count and the union of its children's bounds
//      live in LDS: each arrival folds its box in with LDS float min/max and
//      then adds to the counter, and the second arrival reads the union back.
//      A wave's LDS operations complete in order, so the union is whole once
//      the counter shows both arrivals; no wave waits on another. Nodes that
//      span chunks are marked in the parent codes by the hierarchy kernel
//      (kSpanBit); their child slot is written and the code listed in the
//      chunk's boundary slots.
//   2. k_refit_span<false>: workgroup g takes the boundary codes of chunks
//      [g*G, (g+1)*G) and finishes, with workgroup-scope atomics on global
//      flags, the nodes whose range lies in those chunks; codes of nodes that
//      reach further go to the top list.
//   3. k_refit_span<true>: one workgroup finishes the top list.
//
// Every walk ends at the root within the tree's depth (at most 65 levels for
// 30-bit keys with position tie-breaks); the level bound only guards against
// a corrupt parent array, and reports it through `error`.
constexpr int kMaxLevels = 128;
constexpr int kRefitThreads = 1024;
constexpr uint32_t kRefitGroup = 16;  // chunks per level-2 workgroup
constexpr int kRefitSpanThreads = 256;
constexpr int kRefitTopThreads = 1024;

__device__ __forceinline__ void put_child(NodeRec* np, uint32_t side, float4 b, uint32_t live) {
  np->box[side] = b;
  if (side) np->live.y = live; else np->live.x = live;
}

__device__ __forceinline__ void push_boundary(uint32_t* slots, uint32_t code, uint32_t* error) {
  const uint32_t at = atomicAdd(&slots[0], 1u);
  if (at < kRefitChunkSlots) slots[1 + at] = code;
  else atomicOr(error, 2u);
}

// LDS of a level-1 workgroup, per node of the chunk (index - c0): the union
// of the bounds of the children that arrived, the node's arrival word and
// its own parent code. The arrival word holds the node's far end (node_end,
// as an offset in the chunk) in bits 20..30 and adds (live leaves << 1) | 1
// per arrival. Each arrival stores its bounds and live count in its slot of
// the node's record; the second also stores the record's meta, rebuilt from
// the far end and the arriving child, so that the hierarchy kernel need not
// write a 16-byte part of every record line. (Writing the whole record from
// the second arrival, with both children's bounds in LDS, was slower: 80 KB
// of LDS per workgroup and four 16-byte stores per record, 806 vs 672 us on
// 10M boxes.)
constexpr uint32_t kEndShift = 20;
template <bool GATHER>
__global__ void __launch_bounds__(kRefitThreads) k_refit_chunk(
    float4* __restrict__ leaf_boxes, uint32_t n, uint32_t chunk_log2,
    const uint32_t* __restrict__ parent, NodeRec* nodes, float4* __restrict__ root_box,
    uint32_t* __restrict__ root_live, uint32_t* error, uint32_t* boundary,
    const uint32_t* __restrict__ node_end, const float4* __restrict__ boxes,
    uint32_t* __restrict__ ids, uint32_t off, uint32_t* __restrict__ leaf_of_id) {
  extern __shared__ float4 s_box[];
  const uint32_t C = 1u << chunk_log2;
  uint32_t* s_cnt = reinterpret_cast<uint32_t*>(s_box + C);
  uint32_t* s_par = s_cnt + C;
  const uint32_t c0 = blockIdx.x << chunk_log2;
  const uint32_t c1 = (c0 + C < n ? c0 + C : n) - 1;  // inclusive
  uint32_t* slots = boundary + (uint64_t)blockIdx.x * kRefitChunkStride;
  for (uint32_t j = threadIdx.x; j < C; j += kRefitThreads) {
    const bool internal = c0 + j + 1 < n;  // internal nodes are 0 .. n - 2
    s_box[j] = empty_box();
    // A node spanning chunks has its far end outside the chunk; level 1
    // never completes it, so the masked offset is not used.
    s_cnt[j] = internal ? ((node_end[c0 + j] - c0) & (C - 1)) << kEndShift : 0u;
    s_par[j] = internal ? parent[c0 + j] : 0u;
  }
  __syncthreads();
  for (uint32_t k = c0 + threadIdx.x; k <= c1; k += kRefitThreads) {
    float4 b;
    if (GATHER) {
      const uint32_t local = ids[k];
      if (local >= n) {  // unreachable for a correct sort; keeps reads in bounds
        atomicOr(error, 2u);
        continue;
      }
      const uint32_t id = local + off;
      if (off) ids[k] = id;
      b = boxes[id];
      leaf_boxes[k] = b;
      leaf_of_id[id] = off + k;
    } else {
      b = leaf_boxes[k];
    }
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
    uint32_t self = k;  // position (leaf) or index (internal) of the arriving child
    for (int level = 0;; level++) {
      const uint32_t p = (code & ~kSpanBit) >> 1, side = code & 1u;
      if (level >= kMaxLevels || p >= n - 1) {
        atomicOr(error, 2u);
        break;
      }
      if (code & kSpanBit) {
        // The node spans chunks: level 2 or 3 finishes it.
        put_child(nodes + p, side, b, live);
        push_boundary(slots, code, error);
        break;
      }
      const uint32_t j = p - c0;
      if (j >= C) {  // not in this chunk: a corrupt parent array
        atomicOr(error, 2u);
        break;
      }
      __hip_atomic_fetch_min(&s_box[j].x, b.x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      __hip_atomic_fetch_min(&s_box[j].y, b.y, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      __hip_atomic_fetch_max(&s_box[j].z, b.z, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      __hip_atomic_fetch_max(&s_box[j].w, b.w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      __atomic_signal_fence(__ATOMIC_SEQ_CST);  // the union before the count, in program order
      const uint32_t old = __hip_atomic_fetch_add(&s_cnt[j], (live << 1) | 1u, __ATOMIC_RELAXED,
                                                  __HIP_MEMORY_SCOPE_WORKGROUP);
      __atomic_signal_fence(__ATOMIC_SEQ_CST);
      put_child(nodes + p, side, b, live);
      if (!(old & 1u)) break;  // first arrival: the sibling carries on
      // Second arrival: the union is whole (a wave's LDS operations complete
      // in order, and the sibling's count shows here). The split is the
      // arriving child's position (left side) or the one before it (right).
      const uint32_t gamma = side ? self - 1 : self;
      nodes[p].meta = node_meta(p, c0 + (old >> kEndShift), gamma);
      nodes[p].live.z = gamma;  // the split, read by the wave walks
      live += (old & ((1u << kEndShift) - 1)) >> 1;
      b = s_box[j];
      if (p == 0) {
        root_box[0] = b;
        root_live[0] = live;
        break;
      }
      code = s_par[j];
      self = p;
    }
  }
}

// Levels 2 and 3. Each code names a child slot already filled; the second
// arrival at a node unites both slots and fills the node's slot in its
// parent. TOP = false: workgroup g owns the nodes whose leaf range lies in
// chunks [g*G, (g+1)*G) and lists codes of nodes reaching further in the top
// list; TOP = true: one workgroup takes the top list to the root.
template <bool TOP>
__global__ void __launch_bounds__(TOP ? kRefitTopThreads : kRefitSpanThreads) k_refit_span(
    const uint32_t* __restrict__ parent, NodeRec* nodes, uint32_t n, uint32_t chunk_log2,
    uint32_t n_chunks, float4* __restrict__ root_box, uint32_t* __restrict__ root_live,
    uint32_t* flags, uint32_t* error, uint32_t* boundary) {
  constexpr uint32_t kThreadsHere = TOP ? kRefitTopThreads : kRefitSpanThreads;
  uint32_t* top = boundary + (uint64_t)n_chunks * kRefitChunkStride;  // [0] count, then codes
  const uint64_t top_cap = refit_top_words(n) - 1;
  const uint32_t g0 = blockIdx.x * kRefitGroup;
  const uint32_t lo = g0 << chunk_log2;
  const uint64_t hi64 = ((uint64_t)(g0 + kRefitGroup) << chunk_log2) - 1;
  const uint32_t hi = hi64 < n ? (uint32_t)hi64 : n - 1;
  uint32_t total;
  if (TOP) {
    total = top[0] < top_cap ? top[0] : (uint32_t)top_cap;
  } else {
    const uint32_t groups = n_chunks - g0 < kRefitGroup ? n_chunks - g0 : kRefitGroup;
    total = groups * kRefitChunkSlots;
  }
  for (uint32_t t = threadIdx.x; t < total; t += kThreadsHere) {
    uint32_t code;
    if (TOP) {
      code = top[1 + t];
    } else {
      const uint32_t* slots = boundary + (uint64_t)(g0 + t / kRefitChunkSlots) * kRefitChunkStride;
      const uint32_t i = t % kRefitChunkSlots;
      const uint32_t cnt = slots[0] < kRefitChunkSlots ? slots[0] : kRefitChunkSlots;
      if (i >= cnt) continue;
      code = slots[1 + i];
    }
    for (int level = 0;; level++) {
      const uint32_t p = (code & ~kSpanBit) >> 1;
      if (level >= kMaxLevels || p >= n - 1) {
        atomicOr(error, 2u);
        break;
      }
      NodeRec* np = nodes + p;
      if (!TOP) {
        const uint4 m = np->meta;
        if (m.z < lo || m.w > hi) {  // reaches beyond this group's chunks
          const uint32_t at = atomicAdd(&top[0], 1u);
          if (at < top_cap) top[1 + at] = code;
          else atomicOr(error, 2u);
          break;
        }
      }
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
      put_child(nodes + ((code & ~kSpanBit) >> 1), code & 1u, b, live);
    }
  }
}

// ------------------------------------------------------------- top ----
// The tree's top (TreeView::top): starting from the root's two children,
// every round splits internal entries into their two children, in place (so
// leaf order is kept), the largest first when the 2 * kTopPairs entries
// would overflow, until no entry can be split. One wavefront; each round is
// one dependent load of node records.
constexpr uint32_t kTopEntries = 2 * kTopPairs;
__global__ void __launch_bounds__(kWave) k_tree_top(const NodeRec* __restrict__ nodes,
                                                     NodeRec* __restrict__ top) {
  __shared__ float4 s_box[2][kTopEntries];
  __shared__ uint32_t s_ref[2][kTopEntries], s_first[2][kTopEntries], s_cnt[2][kTopEntries],
      s_live[2][kTopEntries];
  const uint32_t lane = threadIdx.x;
  // The two children of internal node `node` as entries at..at+1 of buffer b.
  auto children = [&](uint32_t node, int b, uint32_t at) {
    const NodeRec nd = nodes[node];
    const uint32_t gamma = nd.meta.x & ~kLeafFlag;
    s_box[b][at] = nd.box[0];
    s_box[b][at + 1] = nd.box[1];
    s_ref[b][at] = nd.meta.x;
    s_ref[b][at + 1] = nd.meta.y;
    s_first[b][at] = nd.meta.z;
    s_first[b][at + 1] = gamma + 1;
    s_cnt[b][at] = gamma - nd.meta.z + 1;
    s_cnt[b][at + 1] = nd.meta.w - gamma;
    s_live[b][at] = nd.live.x;
    s_live[b][at + 1] = nd.live.y;
  };
  if (lane == 0) children(0, 0, 0);
  __syncthreads();
  uint32_t m = 2, cur = 0;  // wave-uniform
  // Each round splits at least one entry or stops, and m <= kTopEntries.
  for (uint32_t round = 0; round < kTopEntries; round++) {
    bool split[2];
    for (int k = 0; k < 2; k++) {
      const uint32_t i = lane + k * kWave;
      split[k] = i < m && !(s_ref[cur][i] & kLeafFlag);
    }
    const uint32_t splittable =
        (uint32_t)__popcll(ballot(split[0])) + (uint32_t)__popcll(ballot(split[1]));
    const uint32_t room = kTopEntries - m;
    if (splittable == 0 || room == 0) break;
    if (splittable > room) {
      // The `room` largest (ties: the earlier entry).
      for (int k = 0; k < 2; k++) {
        if (!split[k]) continue;
        const uint32_t i = lane + k * kWave, c = s_cnt[cur][i];
        uint32_t rank = 0;
        for (uint32_t j = 0; j < m; j++) {
          if (s_ref[cur][j] & kLeafFlag) continue;
          const uint32_t cj = s_cnt[cur][j];
          rank += (cj > c || (cj == c && j < i)) ? 1u : 0u;
        }
        split[k] = rank < room;
      }
    }
    // Entry i moves to i + (splits before i).
    const uint32_t a0 = split[0] ? 1u : 0u, a1 = split[1] ? 1u : 0u;
    const uint32_t inc0 = wave_inclusive_sum(a0);
    const uint32_t tot0 = readlane_u32(inc0, kWave - 1);
    const uint32_t inc1 = wave_inclusive_sum(a1);
    const uint32_t before[2] = {inc0 - a0, tot0 + inc1 - a1};
    const int nxt = cur ^ 1;
    for (int k = 0; k < 2; k++) {
      const uint32_t i = lane + k * kWave;
      if (i >= m) continue;
      const uint32_t at = i + before[k];
      if (split[k]) {
        children(s_ref[cur][i], nxt, at);
      } else {
        s_box[nxt][at] = s_box[cur][i];
        s_ref[nxt][at] = s_ref[cur][i];
        s_first[nxt][at] = s_first[cur][i];
        s_cnt[nxt][at] = s_cnt[cur][i];
        s_live[nxt][at] = s_live[cur][i];
      }
    }
    m += tot0 + readlane_u32(inc1, kWave - 1);
    cur = nxt;
    __syncthreads();
  }
  // Record `lane`: entries 2 lane (left) and 2 lane + 1 (right); entries
  // past m are empty (the right one then covers no leaves).
  const uint32_t i0 = 2 * lane, i1 = i0 + 1;
  const bool ok0 = i0 < m, ok1 = i1 < m;
  const uint32_t last0 = ok0 ? s_first[cur][i0] + s_cnt[cur][i0] - 1 : 0u;
  NodeRec r;
  r.box[0] = ok0 ? s_box[cur][i0] : empty_box();
  r.box[1] = ok1 ? s_box[cur][i1] : empty_box();
  r.meta = make_uint4(ok0 ? s_ref[cur][i0] : kLeafFlag, ok1 ? s_ref[cur][i1] : kLeafFlag,
                      ok0 ? s_first[cur][i0] : 0u,
                      ok1 ? s_first[cur][i1] + s_cnt[cur][i1] - 1 : last0);
  r.live = make_uint4(ok0 ? s_live[cur][i0] : 0u, ok1 ? s_live[cur][i1] : 0u, last0, 0u);
  top[lane] = r;
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
                         uint32_t* node_end, hipStream_t stream) {
  if (n < 2) return hipSuccess;
  hipLaunchKernelGGL(k_karras, dim3(grid_for(n - 1)), dim3(kThreads), 0, stream, sorted_keys, n,
                     nodes, parent, node_end, refit_chunk_log2(n));
  return hipGetLastError();
}

hipError_t launch_refit(const RefitArgs& a, hipStream_t stream) {
  const uint32_t n = a.n;
  if (n == 0) return hipSuccess;
  const uint32_t lg = refit_chunk_log2(n);
  const uint32_t chunks = (uint32_t)(((uint64_t)n + (1u << lg) - 1) >> lg);
  // Level-1 slots (count + codes per chunk) and the top list's count start at zero.
  hipError_t e = launch_fill_u32(a.boundary, 0u, (uint64_t)chunks * kRefitChunkStride + 1, stream);
  if (e != hipSuccess) return e;
  if (chunks > 1) {
    hipLaunchKernelGGL(k_clear_flags, dim3(grid_for(n - 1)), dim3(kThreads), 0, stream, a.flags,
                       n - 1);
  }
  const size_t lds = refit_chunk_lds_bytes(lg);
  if (a.gather)
    hipLaunchKernelGGL(k_refit_chunk<true>, dim3(chunks), dim3(kRefitThreads), lds, stream,
                       a.leaf_boxes, n, lg, a.parent, a.nodes, a.root_box, a.root_live, a.error,
                       a.boundary, a.node_end, a.boxes, a.ids, a.off, a.leaf_of_id);
  else
    hipLaunchKernelGGL(k_refit_chunk<false>, dim3(chunks), dim3(kRefitThreads), lds, stream,
                       a.leaf_boxes, n, lg, a.parent, a.nodes, a.root_box, a.root_live, a.error,
                       a.boundary, a.node_end, nullptr, nullptr, 0u, nullptr);
  if (chunks > 1) {
    const uint32_t groups = (chunks + kRefitGroup - 1) / kRefitGroup;
    hipLaunchKernelGGL(k_refit_span<false>, dim3(groups), dim3(kRefitSpanThreads), 0, stream,
                       a.parent, a.nodes, n, lg, chunks, a.root_box, a.root_live, a.flags, a.error,
                       a.boundary);
    hipLaunchKernelGGL(k_refit_span<true>, dim3(1), dim3(kRefitTopThreads), 0, stream, a.parent,
                       a.nodes, n, lg, chunks, a.root_box, a.root_live, a.flags, a.error,
                       a.boundary);
  }
  if (a.top && n >= kTopMinLeaves)
    hipLaunchKernelGGL(k_tree_top, dim3(1), dim3(kWave), 0, stream, a.nodes, a.top);
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
