// LBVH traversal kernels: point, envelope and line-segment queries with exact
// FP32 (or, on a double index, exact FP64) predicates, replacing the
// reference's OptiX raygen/intersection programs
// (src/shaders/shaders_contains_point_query_2d.cu:14-49,
//  shaders_contains_envelope_query_2d.cu:14-51,
//  shaders_intersects_envelope_query_2d.cu:14-128,
//  shaders_intersects_line_query_2d.cu:14-70).
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
#include "range_log.h"
#include "rtspatial/device/traverse.h"
#include "rtspatial/device/wave.h"

namespace rts {
namespace {

constexpr int kThreads = 256;
constexpr int kWaves = kThreads / kWave;
constexpr int kStageCap = 1024;  // pairs staged per wavefront before a flush

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
    if (slot < cap) {
      const rts_pair pr = stage[j];
      store_pair(out + slot, pr.geom_id, pr.query_id);
    }
  }
  lds_fence();
}

// One query per lane. The wave iterates until all of its lanes are done; in
// each iteration a lane either expands one node or tests one leaf box.
template <int PRED, bool POINTS, bool COLLECT, bool F64>
__global__ void __launch_bounds__(kThreads) k_query_thread(const QueryLaunch q) {
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

  query_t<PRED, F64> qb{};
  if (working) load_q(q.queries, POINTS, qi, qb);

  uint32_t stack[kStackDepth];
  int sp = 0;
  uint32_t cur = 1, end = 0;  // leaf range being scanned: [cur, end]
  if (working) {
    const float4 rb = t.root_box[0];
    if (!node_test<PRED, POINTS>(rb, qb)) {
      working = false;
    } else if (t.n <= t.leaf_size || t.n == 1 || inside_test<PRED, POINTS>(rb, qb)) {
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
          if (node_test<PRED, POINTS>(rbx, qb)) {
            if (m.y & kLeafFlag) {
              lane_push(stack, sp, m.y, t.error);
            } else {
              const uint32_t size = m.w - gamma;
              lane_push(stack, sp,
                        (size <= t.leaf_size || inside_test<PRED, POINTS>(rbx, qb))
                            ? (m.y | kRangeFlag) : m.y,
                        t.error);
            }
          }
          if (node_test<PRED, POINTS>(lb, qb)) {
            if (m.x & kLeafFlag) {
              lane_push(stack, sp, m.x, t.error);
            } else {
              const uint32_t size = gamma - m.z + 1;
              lane_push(stack, sp,
                        (size <= t.leaf_size || inside_test<PRED, POINTS>(lb, qb))
                            ? (m.x | kRangeFlag) : m.x,
                        t.error);
            }
          }
        }
      } else {
        working = false;
      }
      if (have && prim < t.n) hit = leaf_test<PRED, POINTS>(t, prim, t.leaf_boxes[prim], qb);
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
// Per-request wave traversal (traverse.h wave_query): kWaveCount writes the
// query's result count to counts[query]; kWaveCountSink adds it to the
// counting handler's counter; kWaveWrite writes the pairs at queue slot
// base + offsets[query] (offsets from the scan of the counts), so the queue
// holds each query's pairs contiguously, in query order. kWaveStaged
// collects in one pass, for walks whose count pass would cost as much as
// the write (segments: no subtree lies inside a segment): each wavefront
// stages its pairs in LDS and flushes them with one device atomic per
// kStageWave pairs; a query's pairs may then be split over flushes.
enum { kWaveCount = 0, kWaveCountSink = 1, kWaveWrite = 2, kWaveStaged = 3 };
constexpr uint32_t kStageWave = kScanStep;  // one scan step's results always fit

struct StagedEmit {
  rts_pair* stage;  // LDS, kStageWave pairs
  unsigned long long* counter;
  rts_pair* out;
  uint64_t cap;
  uint32_t fill = 0;  // wave-uniform: pairs staged
  int64_t off = 0;    // result k of the current step goes to stage[off + k]
  __device__ void flush() {
    if (fill) {
      lds_fence();
      flush_stage(stage, fill, counter, out, cap);
      fill = 0;
    }
  }
  __device__ void begin_step(uint64_t first_k, uint32_t n) {
    if (fill + n > kStageWave) flush();
    off = (int64_t)fill - (int64_t)first_k;
    fill += n;
  }
  __device__ void operator()(uint64_t k, uint32_t geom_id, uint32_t query_id) const {
    rts_pair pr;
    pr.geom_id = geom_id;
    pr.query_id = query_id;
    stage[(uint32_t)(off + (int64_t)k)] = pr;
  }
};

// One wavefront per workgroup, so that a wave's LDS is released by itself.
// Float box walks are held to 72 VGPRs (7 waves per SIMD), which the tree
// top's first step would otherwise push just past; segment and double walks
// would spill under that bound.
template <int PRED, bool POINTS, int MODE, bool F64>
__global__ void __launch_bounds__(kWave)
__attribute__((amdgpu_waves_per_eu((PRED == kPredLine || F64) ? 1 : 7)))
k_query_wave(const QueryLaunch q, const WaveArgs a) {
  __shared__ WaveLds s_lds;
  __shared__ rts_pair s_stage[MODE == kWaveStaged ? kStageWave : 1];
  const TreeView t = q.tree;
  if (t.n == 0) return;
  StagedEmit se{s_stage, nullptr, nullptr, 0};
  if (MODE == kWaveStaged) {
    const rts_dev_queue* dq = static_cast<const rts_dev_queue*>(q.sink.arg);
    se.counter = dq->counter;
    se.out = dq->data;
    se.cap = dq->capacity;
  }
  uint32_t slice = blockIdx.x % kXcds;
  LogCursor lc = MODE == kWaveCount ? log_cursor(a.log) : LogCursor{};
  // Every iteration takes a query or finds none left: q.n + 1 iterations
  // bound the loop.
  for (uint32_t it = 0; it <= q.n; it++) {
    const uint32_t lin = grab_query(a.tickets, q.n, slice);
    if (lin >= q.n) break;
    const uint32_t qi = q.order ? q.order[lin] : lin;
    query_t<PRED, F64> qb;
    load_q(q.queries, POINTS, qi, qb);
    if (MODE == kWaveStaged) {
      wave_query<PRED, POINTS, true>(t, qb, qi + q.qid_base, se, s_lds);
      continue;
    }
    // Two-pass calls: the count walk logs its leaf ranges (range_log.h), the
    // write pass replays them instead of walking again.
    if (MODE == kWaveWrite) {
      const rts_dev_queue* dq = static_cast<const rts_dev_queue*>(q.sink.arg);
      RangeLog lg = log_for_write(a.log, qi);
      wave_query<PRED, POINTS, true, true>(
          t, qb, qi + q.qid_base, QueueEmit{dq->data, dq->capacity, a.base[0] + a.offsets[qi]},
          s_lds, &lg);
      continue;
    }
    uint64_t all;
    if (MODE == kWaveCount) {
      RangeLog lg = log_for_count(lc);
      all = wave_query<PRED, POINTS, false, true>(t, qb, 0, NoEmit{}, s_lds, &lg);
      log_commit(lc, a.log, qi, lg);
    } else {
      all = wave_query<PRED, POINTS, false>(t, qb, 0, NoEmit{}, s_lds);
    }
    if (lane_id() == 0) {
      if (MODE == kWaveCount) {
        a.counts[qi] = (uint32_t)all;
      } else if (all) {
        atomicAdd(static_cast<unsigned long long*>(const_cast<void*>(q.sink.arg)),
                  (unsigned long long)all);
      }
    }
  }
  if (MODE == kWaveStaged) se.flush();
}

// Persistent grid: 32 single-wave workgroups per CU, the most waves a CU
// holds. Occupancy (VGPRs, LDS) admits fewer at once; the rest start as
// waves finish and exit as soon as they find every slice drained.
uint32_t wave_grid_blocks() {
  static const uint32_t blocks = [] {
    int dev = 0, cus = 0;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    return (uint32_t)((cus > 0 ? cus : 256) * 32);
  }();
  return blocks;
}

template <int PRED, bool POINTS, int MODE, bool F64>
hipError_t launch_wave_mode(const QueryLaunch& q, const WaveArgs& a_in, hipStream_t stream) {
  hipError_t e = launch_fill_u32(a_in.tickets, 0u, kXcds * kTicketStride, stream);
  if (e != hipSuccess) return e;
  const uint32_t resident = wave_grid_blocks();
  uint32_t blocks = resident < q.n ? resident : q.n;
  blocks = (blocks + kXcds - 1) / kXcds * kXcds;  // whole rounds over the XCDs
  WaveArgs a = a_in;
  if (blocks) a.log.per_wave = (uint32_t)(a.log.n_entries / blocks);
  hipLaunchKernelGGL((k_query_wave<PRED, POINTS, MODE, F64>), dim3(blocks), dim3(kWave), 0, stream,
                     q, a);
  return hipGetLastError();
}

template <int PRED, bool POINTS, bool F64>
hipError_t launch_wave_f(const QueryLaunch& q, const WaveArgs& a, int phase, hipStream_t stream) {
  switch (phase) {
    case kWaveCount:
      return launch_wave_mode<PRED, POINTS, kWaveCount, F64>(q, a, stream);
    case kWaveCountSink:
      return launch_wave_mode<PRED, POINTS, kWaveCountSink, F64>(q, a, stream);
    case kWaveStaged:
      return launch_wave_mode<PRED, POINTS, kWaveStaged, F64>(q, a, stream);
    default:
      return launch_wave_mode<PRED, POINTS, kWaveWrite, F64>(q, a, stream);
  }
}

template <int PRED, bool POINTS>
hipError_t launch_wave(const QueryLaunch& q, const WaveArgs& a, int phase, hipStream_t stream) {
  return q.f64 ? launch_wave_f<PRED, POINTS, true>(q, a, phase, stream)
               : launch_wave_f<PRED, POINTS, false>(q, a, phase, stream);
}

template <int PRED, bool POINTS, bool F64>
hipError_t launch_pred_f(const QueryLaunch& q, hipStream_t stream) {
  const dim3 grid((q.n + kThreads - 1) / kThreads), block(kThreads);
  if (q.sink.mode == RTS_HANDLER_COLLECT)
    hipLaunchKernelGGL((k_query_thread<PRED, POINTS, true, F64>), grid, block, 0, stream, q);
  else
    hipLaunchKernelGGL((k_query_thread<PRED, POINTS, false, F64>), grid, block, 0, stream, q);
  return hipGetLastError();
}

template <int PRED, bool POINTS>
hipError_t launch_pred(const QueryLaunch& q, hipStream_t stream) {
  return q.f64 ? launch_pred_f<PRED, POINTS, true>(q, stream)
               : launch_pred_f<PRED, POINTS, false>(q, stream);
}

}  // namespace

hipError_t launch_query(const QueryLaunch& q, hipStream_t stream) {
  if (q.n == 0) return hipSuccess;
  if (q.point_queries) return launch_pred<RTS_CONTAINS, true>(q, stream);
  switch (q.pred) {
    case RTS_INTERSECTS: return launch_pred<RTS_INTERSECTS, false>(q, stream);
    case RTS_CONTAINS: return launch_pred<RTS_CONTAINS, false>(q, stream);
    case RTS_WITHIN: return launch_pred<RTS_WITHIN, false>(q, stream);
    case kPredLine: return launch_pred<kPredLine, false>(q, stream);
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
    case kPredLine: return launch_wave<kPredLine, false>(q, a, phase, stream);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace rts
