// LBVH traversal kernels: point and envelope queries with exact FP32
// predicates, replacing the reference's OptiX raygen/intersection programs
// (src/shaders/shaders_contains_point_query_2d.cu:14-49,
//  shaders_contains_envelope_query_2d.cu:14-51,
//  shaders_intersects_envelope_query_2d.cu:14-128).
//
// Node tests are necessary conditions of the leaf predicate on exact FP32
// bounds, so pruning never drops a true pair; every reported pair passed the
// exact leaf predicate (geom/envelope.cuh:29-69). A child whose bounds lie
// inside the query (intersects / within) is not descended: its leaf range is
// scanned directly, since every live box under it matches.
//
// Output: results of a wavefront are staged in LDS and flushed with one
// device-scope atomic per ~1000 pairs and fully coalesced 8-byte stores, in
// place of the reference's one atomic per result (utils/queue.h:22-27).
#include "rts_internal.h"
#include "wave.h"

namespace rts {
namespace {

constexpr int kThreads = 256;
constexpr int kWaves = kThreads / kWave;
constexpr int kStageCap = 1024;  // pairs staged per wavefront before a flush

template <int PRED, bool POINTS>
struct Pred;

// Point queries: Envelope::Contains(point) (envelope.cuh:29-38).
template <int P>
struct Pred<P, true> {
  __device__ static bool node(const float4 n, const float4 q) { return contains_point(n, q.x, q.y); }
  __device__ static bool inside(const float4, const float4) { return false; }
  __device__ static bool leaf(const float4 b, const float4 q) { return contains_point(b, q.x, q.y); }
};
// Envelope intersects (envelope.cuh:55-69).
template <>
struct Pred<RTS_INTERSECTS, false> {
  __device__ static bool node(const float4 n, const float4 q) { return intersects(n, q); }
  __device__ static bool inside(const float4 n, const float4 q) { return contains_box(q, n); }
  __device__ static bool leaf(const float4 b, const float4 q) { return intersects(b, q); }
};
// Data box contains the query (envelope.cuh:40-53, the reference's kContains).
template <>
struct Pred<RTS_CONTAINS, false> {
  __device__ static bool node(const float4 n, const float4 q) { return contains_box(n, q); }
  __device__ static bool inside(const float4, const float4) { return false; }
  __device__ static bool leaf(const float4 b, const float4 q) { return contains_box(b, q); }
};
// Data box inside the query: the node must meet the query.
template <>
struct Pred<RTS_WITHIN, false> {
  __device__ static bool node(const float4 n, const float4 q) { return intersects(n, q); }
  __device__ static bool inside(const float4 n, const float4 q) { return contains_box(q, n); }
  __device__ static bool leaf(const float4 b, const float4 q) { return contains_box(q, b); }
};

__device__ __forceinline__ void lds_fence() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

// Flushes `fill` staged pairs of this wave to the queue.
__device__ __forceinline__ void flush_stage(const rts_pair* stage, uint32_t fill,
                                            unsigned long long* counter, rts_pair* out,
                                            uint64_t cap) {
  const uint32_t lane = lane_id();
  unsigned long long base = 0;
  if (lane == 0) base = atomicAdd(counter, (unsigned long long)fill);
  base = readlane_u64(base, 0);
  lds_fence();
  for (uint32_t j = lane; j < fill; j += kWave) {
    const uint64_t slot = base + j;
    if (slot < cap) out[slot] = stage[j];
  }
  lds_fence();
}

// One query per lane. The wave iterates until all of its lanes are done; in
// each iteration a lane either expands one node or tests one leaf box.
template <int PRED, bool POINTS, bool COLLECT>
__global__ void __launch_bounds__(kThreads) k_query_thread(const QueryLaunch q) {
  using P = Pred<PRED, POINTS>;
  __shared__ rts_pair s_stage[COLLECT ? kWaves : 1][COLLECT ? kStageCap : 1];

  unsigned long long* counter;
  rts_pair* out = nullptr;
  uint64_t cap = 0;
  if (COLLECT) {
    const rts_dev_queue* dq = static_cast<const rts_dev_queue*>(q.sink.arg);
    out = dq->data;
    cap = dq->capacity;
    counter = dq->counter;
  } else {
    counter = static_cast<unsigned long long*>(const_cast<void*>(q.sink.arg));
  }

  const TreeView t = q.tree;
  const uint32_t wave = threadIdx.x / kWave;
  const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
  bool working = i < q.n && t.n > 0;
  const uint32_t qi = working ? (q.order ? q.order[i] : i) : 0;
  const uint32_t qid = qi + q.qid_base;

  float4 qb = make_float4(0.f, 0.f, 0.f, 0.f);
  if (working) {
    if (POINTS) {
      const float2 p = static_cast<const float2*>(q.queries)[qi];
      qb = make_float4(p.x, p.y, p.x, p.y);
    } else {
      qb = static_cast<const float4*>(q.queries)[qi];
    }
  }

  uint32_t stack[kStackDepth];
  int sp = 0;
  uint32_t cur = 1, end = 0;  // leaf range being scanned: [cur, end]
  if (working) {
    const float4 rb = t.root_box[0];
    if (!P::node(rb, qb)) {
      working = false;
    } else if (t.n <= t.leaf_size || t.n == 1 || P::inside(rb, qb)) {
      cur = 0;
      end = t.n - 1;
    } else {
      stack[sp++] = 0;
    }
  }

  uint64_t my_count = 0;
  uint32_t fill = 0;  // wave-uniform
  rts_pair* stage = s_stage[COLLECT ? wave : 0];
  // A correct tree is walked in at most 2n-1 node expansions plus n leaf
  // tests per query; the bound only stops a walk of a corrupt tree.
  const uint64_t max_steps = 4ull * t.n + 64;
  uint64_t steps = 0;

  while (ballot(working)) {
    bool hit = false;
    uint32_t prim = 0;
    if (working && ++steps > max_steps) {
      atomicOr(t.error, kErrTraversal);
      working = false;
    }
    if (working) {
      bool have = false;
      if (cur <= end) {
        prim = cur++;
        have = true;
      } else if (sp > 0) {
        const uint32_t e = stack[--sp];
        if (e & kLeafFlag) {
          prim = e & ~kLeafFlag;
          have = true;
        } else if (e & kRangeFlag) {
          if ((e & kIndexMask) < t.n - 1) {
            const uint4 m = t.nodes[e & kIndexMask].meta;
            prim = m.z;
            cur = m.z + 1;
            end = m.w < t.n ? m.w : t.n - 1;
            have = true;
          }
        } else if (e < t.n - 1) {
          const NodeRec* nd = t.nodes + e;
          const uint4 m = nd->meta;
          const float4 lb = nd->box[0];
          const float4 rbx = nd->box[1];
          const uint32_t gamma = m.x & ~kLeafFlag;
          // right child first so that the left one is expanded next
          if (P::node(rbx, qb)) {
            if (m.y & kLeafFlag) {
              if (sp < kStackDepth) stack[sp++] = m.y;
            } else {
              const uint32_t size = m.w - gamma;
              if (sp < kStackDepth)
                stack[sp++] = (size <= t.leaf_size || P::inside(rbx, qb)) ? (m.y | kRangeFlag) : m.y;
            }
          }
          if (P::node(lb, qb)) {
            if (m.x & kLeafFlag) {
              if (sp < kStackDepth) stack[sp++] = m.x;
            } else {
              const uint32_t size = gamma - m.z + 1;
              if (sp < kStackDepth)
                stack[sp++] = (size <= t.leaf_size || P::inside(lb, qb)) ? (m.x | kRangeFlag) : m.x;
            }
          }
        }
      } else {
        working = false;
      }
      if (have && prim < t.n) hit = P::leaf(t.leaf_boxes[prim], qb);
    }

    if (COLLECT) {
      const uint64_t mask = ballot(hit);
      if (mask) {
        if (hit) {
          rts_pair pr;
          pr.geom_id = t.leaf_ids[prim];
          pr.query_id = qid;
          stage[fill + mask_rank(mask)] = pr;
        }
        fill += (uint32_t)__popcll(mask);
        if (fill > kStageCap - kWave) {
          lds_fence();
          flush_stage(stage, fill, counter, out, cap);
          fill = 0;
        }
      }
    } else {
      my_count += hit ? 1u : 0u;
    }
  }

  if (COLLECT) {
    if (fill) {
      lds_fence();
      flush_stage(stage, fill, counter, out, cap);
    }
  } else {
    const uint64_t total = wave_sum_u64(my_count);
    if (lane_id() == 0 && total) atomicAdd(counter, (unsigned long long)total);
  }
}

// ---------------------------------------------------------------- wave ----
// One query per wavefront. The wave expands up to 64 stack entries per step
// (one per lane) from a stack in LDS; children that are leaves, small
// subtrees (<= leaf_size leaves) or lie inside the query become leaf ranges
// in an LDS list. Once the list holds enough leaves the wave scans it 64
// leaves at a time -- consecutive lanes read consecutive leaves -- and writes
// the matches of this query contiguously.
//
// Modes: kWaveCount writes the query's result count to counts[query];
// kWaveCountSink adds it to the counting handler's counter; kWaveWrite
// writes the pairs at queue slot base + offsets[query] (offsets from the
// scan of the counts), so the queue holds each query's pairs contiguously,
// in query order.
constexpr int kWaveStack = 1024;    // LDS stack entries per wave
constexpr int kWaveRanges = 256;    // leaf ranges per wave before a scan
constexpr int kFlushLeaves = 512;   // scan once this many leaves are queued
constexpr int kScanUnroll = 4;      // leaves per lane per scan step
constexpr uint32_t kScanStep = kScanUnroll * kWave;
enum { kWaveCount = 0, kWaveCountSink = 1, kWaveWrite = 2 };

// Workgroups are dealt to the 8 XCDs round-robin; remapping the launch index
// gives each XCD (and its L2) one contiguous slice of the query order.
constexpr uint32_t kXcds = 8;
__device__ __forceinline__ uint32_t xcd_slice_block(uint32_t b, uint32_t grid) {
  const uint32_t per = grid / kXcds;  // grid is a multiple of kXcds
  return (b % kXcds) * per + b / kXcds;
}

// Range entries: first leaf | kAllMatch when every leaf of the range is a
// live box inside the query (no test needed).
constexpr uint32_t kAllMatch = 0x80000000u;

template <int PRED, bool POINTS, int MODE>
__global__ void __launch_bounds__(kThreads) k_query_wave(const QueryLaunch q, const WaveArgs a) {
  using P = Pred<PRED, POINTS>;
  __shared__ uint32_t s_stack[kWaves][kWaveStack];
  __shared__ uint32_t s_rfirst[kWaves][kWaveRanges];
  __shared__ uint32_t s_rpre[kWaves][kWaveRanges + 1];

  const TreeView t = q.tree;
  const uint32_t w = threadIdx.x / kWave, lane = lane_id();
  const uint32_t lin = xcd_slice_block(blockIdx.x, gridDim.x) * kWaves + w;
  if (lin >= q.n || t.n == 0) return;  // wave-uniform
  const uint32_t qi = q.order ? q.order[lin] : lin;
  const uint32_t qid = qi + q.qid_base;
  float4 qb;
  if (POINTS) {
    const float2 p = static_cast<const float2*>(q.queries)[qi];
    qb = make_float4(p.x, p.y, p.x, p.y);
  } else {
    qb = static_cast<const float4*>(q.queries)[qi];
  }

  rts_pair* out = nullptr;
  uint64_t cap = 0, obase = 0;
  if (MODE == kWaveWrite) {
    const rts_dev_queue* dq = static_cast<const rts_dev_queue*>(q.sink.arg);
    out = dq->data;
    cap = dq->capacity;
    obase = a.base[0] + a.offsets[qi];
  }
  constexpr bool kCounting = MODE != kWaveWrite;

  uint32_t* stack = s_stack[w];
  uint32_t* rfirst = s_rfirst[w];
  uint32_t* rpre = s_rpre[w];
  uint32_t sp = 0, m = 0, total = 0;  // wave-uniform
  uint64_t emitted = 0;               // wave-uniform: results of scanned ranges
  uint64_t inside_live = 0;           // per lane: counting mode, subtrees inside the query

  const float4 rb = t.root_box[0];
  if (P::node(rb, qb)) {
    const bool inside = P::inside(rb, qb);
    if (kCounting && inside) {
      if (lane == 0) inside_live = t.root_live[0];
    } else if (t.n <= t.leaf_size || t.n == 1 || inside) {
      if (lane == 0) {
        rfirst[0] = (inside && t.root_live[0] == t.n) ? kAllMatch : 0u;
        rpre[0] = 0;
        rpre[1] = t.n;
      }
      m = 1;
      total = t.n;
    } else {
      if (lane == 0) stack[0] = 0;
      sp = 1;
    }
  }
  lds_fence();

  const uint64_t max_iters = 4ull * t.n + 64;
  uint64_t iters = 0;
  while (sp > 0 || m > 0) {
    if (++iters > max_iters) {
      if (lane == 0) atomicOr(t.error, kErrTraversal);
      break;
    }
    if (sp > 0 && total < kFlushLeaves && m + 2 * kWave <= kWaveRanges) {
      // ---- expand up to 64 stack entries, one per lane
      const uint32_t c = sp < (uint32_t)kWave ? sp : (uint32_t)kWave;
      const bool act = lane < c;
      uint32_t e = act ? stack[sp - 1 - lane] : 0u;
      lds_fence();
      sp -= c;
      // Near a full stack, subtrees are scanned as ranges instead of pushed.
      const bool degrade = sp + 2 * kWave > (uint32_t)kWaveStack;
      uint32_t push0 = 0, push1 = 0, np = 0;
      uint32_t rf0 = 0, rn0 = 0, rf1 = 0, rn1 = 0, nr = 0;
      if (act && e < t.n - 1) {
        const NodeRec* nd = t.nodes + e;
        const uint4 mt = nd->meta;
        const uint4 lv = nd->live;
        const float4 cb[2] = {nd->box[0], nd->box[1]};
        const uint32_t live[2] = {lv.x, lv.y};
        const uint32_t gamma = mt.x & ~kLeafFlag;
        const uint32_t ref[2] = {mt.x, mt.y};
        const uint32_t first[2] = {mt.z, gamma + 1};
        const uint32_t cnt[2] = {gamma - mt.z + 1, mt.w - gamma};
#pragma unroll
        for (int side = 0; side < 2; side++) {
          if (!P::node(cb[side], qb)) continue;
          const bool inside = P::inside(cb[side], qb);
          if (kCounting && inside) {
            inside_live += live[side];
          } else if ((ref[side] & kLeafFlag) || cnt[side] <= t.leaf_size || degrade || inside) {
            uint32_t f = first[side];
            if (inside && live[side] == cnt[side]) f |= kAllMatch;
            if (nr == 0) {
              rf0 = f;
              rn0 = cnt[side];
            } else {
              rf1 = f;
              rn1 = cnt[side];
            }
            nr++;
          } else {
            if (np == 0) push0 = ref[side]; else push1 = ref[side];
            np++;
          }
        }
      }
      const uint32_t pinc = wave_inclusive_sum(np);
      const uint32_t ptot = readlane_u32(pinc, kWave - 1);
      const uint32_t pex = pinc - np;
      if (np > 0) stack[sp + pex] = push0;
      if (np > 1) stack[sp + pex + 1] = push1;
      sp += ptot;
      const uint32_t rinc = wave_inclusive_sum(nr);
      const uint32_t rtot = readlane_u32(rinc, kWave - 1);
      const uint32_t rex = rinc - nr;
      const uint32_t sz = rn0 + rn1;
      const uint32_t sinc = wave_inclusive_sum(sz);
      const uint32_t stot = readlane_u32(sinc, kWave - 1);
      const uint32_t pre = total + (sinc - sz);
      if (nr > 0) {
        rfirst[m + rex] = rf0;
        rpre[m + rex] = pre;
      }
      if (nr > 1) {
        rfirst[m + rex + 1] = rf1;
        rpre[m + rex + 1] = pre + rn0;
      }
      m += rtot;
      total += stot;
      if (lane == 0) rpre[m] = total;
      lds_fence();
    } else {
      // ---- scan the queued leaf ranges, kScanUnroll x 64 leaves per step:
      // all loads of a step are issued before any result is written.
      uint32_t e0 = 0;
      for (uint32_t base = 0; base < total; base += kScanStep) {
        uint32_t leaf[kScanUnroll];
        bool hit[kScanUnroll];
        // Fast path: the whole step lies in entry e0.
        const bool one_range = rpre[e0 + 1] >= base + kScanStep;
        uint32_t lo = e0;
#pragma unroll
        for (int u = 0; u < kScanUnroll; u++) {
          const uint32_t j = base + u * kWave + lane;
          hit[u] = false;
          leaf[u] = 0;
          if (j < total) {
            if (!one_range) {
              uint32_t hi = m - 1;
              while (lo < hi) {
                const uint32_t mid = (lo + hi + 1) >> 1;
                if (rpre[mid] <= j) lo = mid; else hi = mid - 1;
              }
            }
            const uint32_t f = rfirst[lo];
            leaf[u] = (f & ~kAllMatch) + (j - rpre[lo]);
            hit[u] = (f & kAllMatch) != 0;
            if (!hit[u] && leaf[u] < t.n) hit[u] = P::leaf(t.leaf_boxes[leaf[u]], qb);
          } else {
            lo = m - 1;
          }
        }
        e0 = readlane_u32(lo, kWave - 1);
        if (MODE == kWaveWrite) {
          uint32_t gid[kScanUnroll];
#pragma unroll
          for (int u = 0; u < kScanUnroll; u++)
            gid[u] = (hit[u] && leaf[u] < t.n) ? t.leaf_ids[leaf[u]] : 0u;
#pragma unroll
          for (int u = 0; u < kScanUnroll; u++) {
            const uint64_t mask = ballot(hit[u]);
            if (hit[u]) {
              const uint64_t slot = obase + emitted + mask_rank(mask);
              if (slot < cap) {
                rts_pair pr;
                pr.geom_id = gid[u];
                pr.query_id = qid;
                out[slot] = pr;
              }
            }
            emitted += (uint64_t)__popcll(mask);
          }
        } else {
#pragma unroll
          for (int u = 0; u < kScanUnroll; u++) emitted += (uint64_t)__popcll(ballot(hit[u]));
        }
      }
      m = 0;
      total = 0;
      lds_fence();
    }
  }

  if (kCounting) {
    const uint64_t all = emitted + wave_sum_u64(inside_live);
    if (lane == 0) {
      if (MODE == kWaveCount) {
        a.counts[qi] = (uint32_t)all;
      } else if (all) {
        atomicAdd(static_cast<unsigned long long*>(const_cast<void*>(q.sink.arg)),
                  (unsigned long long)all);
      }
    }
  }
}

template <int PRED, bool POINTS>
hipError_t launch_wave(const QueryLaunch& q, const WaveArgs& a, int phase, hipStream_t stream) {
  uint32_t blocks = (q.n + kWaves - 1) / kWaves;
  blocks = (blocks + kXcds - 1) / kXcds * kXcds;  // whole rounds over the XCDs
  const dim3 grid(blocks), block(kThreads);
  switch (phase) {
    case kWaveCount:
      hipLaunchKernelGGL((k_query_wave<PRED, POINTS, kWaveCount>), grid, block, 0, stream, q, a);
      break;
    case kWaveCountSink:
      hipLaunchKernelGGL((k_query_wave<PRED, POINTS, kWaveCountSink>), grid, block, 0, stream, q, a);
      break;
    default:
      hipLaunchKernelGGL((k_query_wave<PRED, POINTS, kWaveWrite>), grid, block, 0, stream, q, a);
  }
  return hipGetLastError();
}

template <int PRED, bool POINTS>
hipError_t launch_pred(const QueryLaunch& q, hipStream_t stream) {
  const dim3 grid((q.n + kThreads - 1) / kThreads), block(kThreads);
  if (q.sink.mode == RTS_HANDLER_COLLECT)
    hipLaunchKernelGGL((k_query_thread<PRED, POINTS, true>), grid, block, 0, stream, q);
  else
    hipLaunchKernelGGL((k_query_thread<PRED, POINTS, false>), grid, block, 0, stream, q);
  return hipGetLastError();
}

}  // namespace

hipError_t launch_query(const QueryLaunch& q, hipStream_t stream) {
  if (q.n == 0) return hipSuccess;
  if (q.point_queries) return launch_pred<RTS_CONTAINS, true>(q, stream);
  switch (q.pred) {
    case RTS_INTERSECTS: return launch_pred<RTS_INTERSECTS, false>(q, stream);
    case RTS_CONTAINS: return launch_pred<RTS_CONTAINS, false>(q, stream);
    case RTS_WITHIN: return launch_pred<RTS_WITHIN, false>(q, stream);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_query_wave(const QueryLaunch& q, const WaveArgs& a, int phase,
                             hipStream_t stream) {
  if (q.n == 0) return hipSuccess;
  if (q.point_queries) return launch_wave<RTS_CONTAINS, true>(q, a, phase, stream);
  switch (q.pred) {
    case RTS_INTERSECTS: return launch_wave<RTS_INTERSECTS, false>(q, a, phase, stream);
    case RTS_CONTAINS: return launch_wave<RTS_CONTAINS, false>(q, a, phase, stream);
    case RTS_WITHIN: return launch_wave<RTS_WITHIN, false>(q, a, phase, stream);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace rts
