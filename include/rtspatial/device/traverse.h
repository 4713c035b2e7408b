// Device-side LBVH traversal cores shared by the library's query kernels
// (one query per lane or per wavefront, per request or in persistent
// batches) and by kernels a caller compiles around its own result handler
// (rtspatial/device/query.h). Results leave through an emitter:
// emit(k, geom_id, query_id) is called for the k-th result of a query.
//
// Node tests are necessary conditions of the leaf predicate on exact FP32
// bounds, so pruning never drops a true pair; every reported pair passed the
// exact leaf predicate (reference geom/envelope.cuh:29-69), except that a
// subtree lying inside the query whose leaves are all live is emitted (or
// counted) without testing: refit bounds cover every live box below a node,
// so each of them matches.
//
// Double indexes (QueryF64) run the same walks on the outward-rounded float
// bounds. Node and "inside" tests compare float bounds with the query rounded
// in the direction that makes each comparison exact (round.h), so they are the
// exact tests of the float bounds against the double query. A leaf is decided
// in FP32 when its float box settles it (necessary test fails, or the box
// shrunk by one ulp -- which lies inside the double box -- passes the
// sufficient test); otherwise, when a coordinate lies within one float ulp of
// the query's, the double box is read and tested exactly.
//
// Line-segment queries (QueryLine, predicate kPredLine) run the same walks:
// a node is entered when the segment may meet its bounds (segment.h, a
// necessary condition), and a leaf is reported when the closed segment meets
// the closed box, decided exactly (on a double index: after the necessary
// test on the outward float box, on the double box).
#pragma once

#include <cstddef>
#include <type_traits>
#include <utility>

#include "rtspatial/device/round.h"
#include "rtspatial/device/segment.h"
#include "rtspatial/device/tree.h"
#include "rtspatial/device/wave.h"

namespace rts {

// A float query: (x0, y0, x1, y1); a point is (x, y, x, y).
struct QueryF32 {
  float4 b;
};
// A double query: its bounds rounded inward and outward, and its coordinates
// (x, y or x0, y0, x1, y1) for the exact test.
struct QueryF64 {
  float4 in;   // (up x0, up y0, down x1, down y1)
  float4 out;  // (down x0, down y0, up x1, up y1)
  const double* q;
};

// A segment query (x1, y1, x2, y2): its coordinates as doubles for the exact
// tests and its bounding box in floats for a first test of a node (rounded
// outward on a double index, so the test stays a necessary condition).
template <bool F64>
struct QueryLine;
template <>
struct QueryLine<false> {
  SegmentF g;
  float4 bb;
};
template <>
struct QueryLine<true> {
  Segment g;
  float4 bb;
};

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

// Double queries (envelope.cuh:29-69 with COORD_T = double). `in` decides
// "float bound <= q.x1" and "float bound >= q.x0" exactly; `out` decides
// "float bound <= q.x0" and "float bound >= q.x1" exactly.
__device__ __forceinline__ float4 shrink_ulp(const float4 b) {
  return make_float4(f32_up(b.x), f32_up(b.y), f32_down(b.z), f32_down(b.w));
}
template <int PRED, bool POINTS>
struct Pred64;
template <int P>
struct Pred64<P, true> {
  __device__ static bool node(const float4 n, const QueryF64& q) { return intersects(n, q.in); }
  __device__ static bool inside(const float4, const QueryF64&) { return false; }
  __device__ static bool nec(const float4 b, const QueryF64& q) { return intersects(b, q.in); }
  __device__ static bool suff(const float4 b, const QueryF64& q) {
    return intersects(shrink_ulp(b), q.in);
  }
  __device__ static bool exact(const Box64& d, const double* q) {
    return q[0] >= d.x0 && q[0] <= d.x1 && q[1] >= d.y0 && q[1] <= d.y1;
  }
};
template <>
struct Pred64<RTS_INTERSECTS, false> {
  __device__ static bool node(const float4 n, const QueryF64& q) { return intersects(n, q.in); }
  __device__ static bool inside(const float4 n, const QueryF64& q) { return contains_box(q.in, n); }
  __device__ static bool nec(const float4 b, const QueryF64& q) { return intersects(b, q.in); }
  __device__ static bool suff(const float4 b, const QueryF64& q) {
    return intersects(shrink_ulp(b), q.in);
  }
  __device__ static bool exact(const Box64& d, const double* q) {
    return q[0] <= d.x1 && q[2] >= d.x0 && q[1] <= d.y1 && q[3] >= d.y0;
  }
};
template <>
struct Pred64<RTS_CONTAINS, false> {
  __device__ static bool node(const float4 n, const QueryF64& q) { return contains_box(n, q.out); }
  __device__ static bool inside(const float4, const QueryF64&) { return false; }
  __device__ static bool nec(const float4 b, const QueryF64& q) { return contains_box(b, q.out); }
  __device__ static bool suff(const float4 b, const QueryF64& q) {
    return contains_box(shrink_ulp(b), q.out);
  }
  __device__ static bool exact(const Box64& d, const double* q) {
    return q[0] >= d.x0 && q[2] <= d.x1 && q[1] >= d.y0 && q[3] <= d.y1;
  }
};
template <>
struct Pred64<RTS_WITHIN, false> {
  __device__ static bool node(const float4 n, const QueryF64& q) { return intersects(n, q.in); }
  __device__ static bool inside(const float4 n, const QueryF64& q) { return contains_box(q.in, n); }
  __device__ static bool nec(const float4 b, const QueryF64& q) { return contains_box(q.out, b); }
  __device__ static bool suff(const float4 b, const QueryF64& q) { return contains_box(q.in, b); }
  __device__ static bool exact(const Box64& d, const double* q) {
    return d.x0 >= q[0] && d.x1 <= q[2] && d.y0 >= q[1] && d.y1 <= q[3];
  }
};

// The tests the walks call, by query type.
template <int PRED, bool POINTS>
__device__ __forceinline__ bool node_test(const float4 n, const QueryF32& q) {
  return Pred<PRED, POINTS>::node(n, q.b);
}
template <int PRED, bool POINTS>
__device__ __forceinline__ bool node_test(const float4 n, const QueryF64& q) {
  return Pred64<PRED, POINTS>::node(n, q);
}
template <int PRED, bool POINTS>
__device__ __forceinline__ bool inside_test(const float4 n, const QueryF32& q) {
  return Pred<PRED, POINTS>::inside(n, q.b);
}
template <int PRED, bool POINTS>
__device__ __forceinline__ bool inside_test(const float4 n, const QueryF64& q) {
  return Pred64<PRED, POINTS>::inside(n, q);
}
// Leaf at sorted position `leaf` with float box b.
template <int PRED, bool POINTS>
__device__ __forceinline__ bool leaf_test(const TreeView&, uint32_t, const float4 b,
                                          const QueryF32& q) {
  return Pred<PRED, POINTS>::leaf(b, q.b);
}
// A float query whose predicate is chosen at run time: RTS_INTERSECTS,
// RTS_WITHIN or RTS_CONTAINS (data box contains the query). A point is the
// degenerate envelope (x, y, x, y) under RTS_CONTAINS: contains_box(b, (x, y,
// x, y)) is contains_point(b, x, y) comparison for comparison, and no subtree
// lies inside a point. The persistent batch kernel answers every request kind
// with one instantiation of each walk this way; `pred` is the same for every
// lane of a wave (a request's), so the tests branch on scalar registers.
struct QueryRT {
  float4 b;
  int pred;
};
template <int PRED, bool POINTS>
__device__ __forceinline__ bool node_test(const float4 n, const QueryRT& q) {
  return q.pred == RTS_CONTAINS ? contains_box(n, q.b) : intersects(n, q.b);
}
template <int PRED, bool POINTS>
__device__ __forceinline__ bool inside_test(const float4 n, const QueryRT& q) {
  return q.pred != RTS_CONTAINS && contains_box(q.b, n);
}
template <int PRED, bool POINTS>
__device__ __forceinline__ bool leaf_test(const TreeView&, uint32_t, const float4 b,
                                          const QueryRT& q) {
  if (q.pred == RTS_INTERSECTS) return intersects(b, q.b);
  if (q.pred == RTS_WITHIN) return contains_box(q.b, b);
  return contains_box(b, q.b);
}
template <int PRED, bool POINTS>
__device__ __forceinline__ bool leaf_test(const TreeView& t, uint32_t leaf, const float4 b,
                                          const QueryF64& q) {
  using P = Pred64<PRED, POINTS>;
  if (!P::nec(b, q)) return false;
  if (P::suff(b, q)) return true;
  return P::exact(t.boxes64[t.leaf_ids[leaf]], q.q);
}
// Segments: no subtree lies "inside" a segment. The float bounding-box test
// comes first; on a float index it is the exact bounding-box test (bb is the
// segment's exact bounding box), on a double index a necessary one.
template <int PRED, bool POINTS>
__device__ __forceinline__ bool node_test(const float4 n, const QueryLine<false>& q) {
  return intersects(n, q.bb) && segment_line_may_cross_f(q.g, n.x, n.y, n.z, n.w);
}
template <int PRED, bool POINTS>
__device__ __forceinline__ bool node_test(const float4 n, const QueryLine<true>& q) {
  return intersects(n, q.bb) && segment_line_may_cross(q.g, n.x, n.y, n.z, n.w);
}
template <int PRED, bool POINTS, bool F64>
__device__ __forceinline__ bool inside_test(const float4, const QueryLine<F64>&) {
  return false;
}
template <int PRED, bool POINTS>
__device__ __forceinline__ bool leaf_test(const TreeView&, uint32_t, const float4 b,
                                          const QueryLine<false>& q) {
  return intersects(b, q.bb) && segment_line_crosses_f(q.g, b.x, b.y, b.z, b.w);
}
template <int PRED, bool POINTS>
__device__ __forceinline__ bool leaf_test(const TreeView& t, uint32_t leaf, const float4 b,
                                          const QueryLine<true>& q) {
  if (!intersects(b, q.bb)) return false;
  if (!segment_line_may_cross(q.g, b.x, b.y, b.z, b.w)) return false;
  const Box64 d = t.boxes64[t.leaf_ids[leaf]];
  return segment_meets_box(q.g, d.x0, d.y0, d.x1, d.y1);
}

// A result pair stored with a non-temporal hint: the walk never reads its
// output back, and streaming it past the caches keeps the node, leaf-box and
// id lines the walk re-reads resident (on the 10M x 1M intersects join the
// write pass took 41.3 ms with ordinary stores, 30.2 ms with these).
static_assert(sizeof(rts_pair) == 8 && offsetof(rts_pair, geom_id) == 0, "pair layout");
__device__ __forceinline__ void store_pair(rts_pair* p, uint32_t geom_id, uint32_t query_id) {
  const uint64_t bits = ((uint64_t)query_id << 32) | geom_id;  // rts_pair: geom_id first
  __builtin_nontemporal_store(bits, reinterpret_cast<uint64_t*>(p));
}

// Emitters. The result queue: the query's k-th pair goes to slot base + k,
// dropped at or past `cap`.
struct QueueEmit {
  rts_pair* out;
  uint64_t cap;
  uint64_t base;
  __device__ void operator()(uint64_t k, uint32_t geom_id, uint32_t query_id) const {
    const uint64_t slot = base + k;
    if (slot < cap) store_pair(out + slot, geom_id, query_id);
  }
};
// Counting walks emit nothing.
struct NoEmit {
  __device__ void operator()(uint64_t, uint32_t, uint32_t) const {}
};

__device__ __forceinline__ void lds_fence() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

__device__ __forceinline__ float4 load_query(const void* queries, bool points, uint32_t i) {
  if (points) {
    const float2 p = static_cast<const float2*>(queries)[i];
    return make_float4(p.x, p.y, p.x, p.y);
  }
  return static_cast<const float4*>(queries)[i];
}

// The query type of a walk: float or double, point / envelope or segment.
template <bool F64, bool LINE = false>
struct QueryOf {
  using type = QueryF32;
};
template <>
struct QueryOf<true, false> {
  using type = QueryF64;
};
template <bool F64>
struct QueryOf<F64, true> {
  using type = QueryLine<F64>;
};
template <int PRED, bool F64>
using query_t = typename QueryOf<F64, PRED == kPredLine>::type;
__device__ __forceinline__ void load_q(const void* queries, bool points, uint32_t i, QueryF32& q) {
  q.b = load_query(queries, points, i);
}
__device__ __forceinline__ void load_q(const void* queries, bool points, uint32_t i, QueryF64& q) {
  const double* p = static_cast<const double*>(queries) + (points ? 2ull : 4ull) * i;
  const double x0 = p[0], y0 = p[1];
  const double x1 = points ? x0 : p[2], y1 = points ? y0 : p[3];
  q.in = make_float4(f32_ru(x0), f32_ru(y0), f32_rd(x1), f32_rd(y1));
  q.out = make_float4(f32_rd(x0), f32_rd(y0), f32_ru(x1), f32_ru(y1));
  q.q = p;
}
// Segment i of a float (x1, y1, x2, y2) array, or of a double one. A segment
// with a NaN coordinate gets a NaN bounding box and so meets nothing.
__device__ __forceinline__ void load_q(const void* queries, bool, uint32_t i, QueryLine<false>& q) {
  const float4 f = static_cast<const float4*>(queries)[i];
  q.g = make_segment_f(f.x, f.y, f.z, f.w);
  q.bb = make_float4(fminf(f.x, f.z), fminf(f.y, f.w), fmaxf(f.x, f.z), fmaxf(f.y, f.w));
  if (!(f.x == f.x && f.y == f.y && f.z == f.z && f.w == f.w))
    q.bb = make_float4(__builtin_nanf(""), 0.0f, 0.0f, 0.0f);
}
__device__ __forceinline__ void load_q(const void* queries, bool, uint32_t i, QueryLine<true>& q) {
  const double* p = static_cast<const double*>(queries) + 4ull * i;
  q.g = make_segment(p[0], p[1], p[2], p[3]);
  const double* s = q.g.s;
  q.bb = make_float4(f32_rd(fmin(s[0], s[2])), f32_rd(fmin(s[1], s[3])), f32_ru(fmax(s[0], s[2])),
                     f32_ru(fmax(s[1], s[3])));
  if (!(s[0] == s[0] && s[1] == s[1] && s[2] == s[2] && s[3] == s[3]))
    q.bb = make_float4(__builtin_nanf(""), 0.0f, 0.0f, 0.0f);
}

// ---------------------------------------------------------------- wave ----
// One query per wavefront. The wave expands up to 64 stack entries per step
// (one per lane) from a stack in LDS; children that are leaves, small
// subtrees (<= leaf_size leaves) or lie inside the query become leaf ranges
// in an LDS list. Once the list holds enough leaves the wave scans it, 512
// leaves per step, and writes the matches of this query contiguously from
// slot `obase`.
constexpr int kWaveStack = 256;     // LDS stack entries per wave
// Wave stack entries are internal node indices (below 2^30); kTopRef + i
// stands for record i of the tree's top.
constexpr uint32_t kTopRef = 1u << 30;
constexpr int kWaveRanges = 256;    // leaf ranges per wave before a scan
// Scan once this many leaves are queued. Deferring the scan past one step's
// 512 leaves leaves fewer part-filled final steps: on the 10M x 1M
// intersects join 44.5 -> 42.4 ms per step (512 -> 2048; 4096 and 8192 no
// better), segment join 12.81 -> 12.63 ms.
constexpr int kFlushLeaves = 2048;
constexpr int kPerLane = 8;         // leaves per lane per scan step
constexpr uint32_t kScanStep = kPerLane * kWave;

// Range entries: first leaf | kAllMatch when every leaf of the range is a
// live box inside the query (no test needed).
constexpr uint32_t kAllMatch = 0x80000000u;
// Leaf indices and flat positions are below 2^30, so a leaf offset (first
// leaf - flat start) biased by 2^30 lies in (0, 2^31) and leaves bit 31 to
// kAllMatch.
constexpr uint32_t kDeltaBias = 1u << 30;

struct WaveLds {
  uint32_t stack[kWaveStack];
  uint32_t rfirst[kWaveRanges];
  uint32_t rpre[kWaveRanges + 1];
  // Scan step map, position p = leaf base + p: the (entry + 1) of the range
  // that starts at p, 0 elsewhere (at most kWaveRanges + 1, so 16 bits).
  uint16_t map[kScanStep];
};

// Returns the query's number of results (wave-uniform). With WRITE every
// pair (geom_id, qid) is emitted, numbered from 0 in emission order; without
// WRITE, subtrees inside the query are counted from their live-leaf counts.
// An emitter may also define begin_step(first_k, n), called by the whole wave
// before the n results numbered first_k.. of one scan step are emitted (an
// emitter that stages results per wave makes room for them there).
template <typename E, typename = void>
struct HasBeginStep : std::false_type {};
template <typename E>
struct HasBeginStep<E, std::void_t<decltype(std::declval<E&>().begin_step(0ull, 0u))>>
    : std::true_type {};

// A walk's leaf-range log (LOG walks). A counting walk records in `entries`
// (room for `cap`) every leaf range that a writing walk of the same query
// scans -- {first leaf | kAllMatch, leaves}, a subtree inside the query
// included -- and sets `logged` to their number (more than cap: they did not
// all fit). A writing walk given n_in <= cap entries scans those ranges
// instead of walking the tree: the same leaves, tested the same way, hence
// the same results (in another order). The write pass of a two-pass query
// then skips the dependent chain of node loads of a second walk.
constexpr uint32_t kLogOverflow = 0xFFFFFFFFu;
// Log entries bypass the caches both ways (written once, read once), so
// that they do not evict the nodes and leaf data the walks re-read.
__device__ __forceinline__ void log_store(uint2* p, uint32_t first, uint32_t leaves) {
  __builtin_nontemporal_store(((uint64_t)leaves << 32) | first, reinterpret_cast<uint64_t*>(p));
}
__device__ __forceinline__ uint2 log_load(const uint2* p) {
  const uint64_t v = __builtin_nontemporal_load(reinterpret_cast<const uint64_t*>(p));
  return make_uint2((uint32_t)v, (uint32_t)(v >> 32));
}
struct RangeLog {
  uint2* entries = nullptr;
  uint32_t cap = 0;
  uint32_t logged = 0;           // counting walk: entries found (wave-uniform)
  uint32_t n_in = kLogOverflow;  // writing walk: entries to scan (kLogOverflow: walk the tree)
};

// A counting walk with BOUND counts every leaf range that needs tests as a
// whole instead of testing its leaves: the result is an upper bound on the
// query's results, found without loading a leaf box (about half the
// counting walk's time on range queries). It logs the same ranges, so a
// writing walk that replays them emits the exact results.
template <int PRED, bool POINTS, bool WRITE, bool LOG = false, bool BOUND = false, typename Q,
          typename Emit>
__device__ uint64_t wave_query(const TreeView& t, const Q& qb, const uint32_t qid,
                               Emit&& emit, WaveLds& L, RangeLog* log = nullptr) {
  constexpr bool kCounting = !WRITE;
  constexpr bool kRecord = LOG && !WRITE;
  constexpr bool kBound = BOUND && !WRITE;
  // Segment walks start at the root: the top's first step would cost their
  // larger walk registers an occupancy step.
  constexpr bool kUseTop = PRED != kPredLine;
  const uint32_t lane = lane_id();
  uint32_t* stack = L.stack;
  uint32_t* rfirst = L.rfirst;
  uint32_t* rpre = L.rpre;
  uint32_t sp = 0, m = 0, total = 0;  // wave-uniform
  uint64_t emitted = 0;               // wave-uniform: results of scanned ranges
  uint64_t inside_live = 0;           // per lane: counting mode, subtrees inside the query
  uint32_t nlog = 0;                  // wave-uniform: kRecord, ranges logged
  uint32_t lpos = 0, lend = 0;        // wave-uniform: replayed log entries [lpos, lend)

  if (LOG && WRITE && log->n_in <= log->cap) {
    lend = log->n_in;  // replay: no walk
  } else {
    const float4 rb = t.root_box[0];
    if (node_test<PRED, POINTS>(rb, qb)) {
      const bool inside = inside_test<PRED, POINTS>(rb, qb);
      if (kCounting && inside) {
        if (lane == 0) inside_live = t.root_live[0];
        if (kRecord) {
          if (lane == 0 && log->cap > 0)
            log_store(log->entries, t.root_live[0] == t.n ? kAllMatch : 0u, t.n);
          nlog = 1;
        }
      } else if (t.n <= t.leaf_size || t.n == 1 || inside) {
        const uint32_t f = (inside && t.root_live[0] == t.n) ? kAllMatch : 0u;
        if (kBound) {
          if (lane == 0) inside_live = t.n;
        } else {
          if (lane == 0) {
            rfirst[0] = f;
            rpre[0] = 0;
            rpre[1] = t.n;
          }
          m = 1;
          total = t.n;
        }
        if (kRecord) {
          if (lane == 0 && log->cap > 0) log_store(log->entries, f, t.n);
          nlog = 1;
        }
      } else if (kUseTop && t.top) {
        // The first step expands the tree's top: one top record per lane.
        stack[lane] = kTopRef + lane;
        sp = kTopPairs;
      } else {
        if (lane == 0) stack[0] = 0;
        sp = 1;
      }
    }
  }
  lds_fence();

  const uint64_t max_iters = 4ull * t.n + 64;
  uint64_t iters = 0;
  while (sp > 0 || m > 0 || lpos < lend) {
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
      uint32_t lf0 = 0, ln0 = 0, lf1 = 0, ln1 = 0, nl = 0;  // kRecord: this lane's log entries
      if (act && (e < t.n - 1 || (kUseTop && e >= kTopRef))) {
        const NodeRec* nd = (kUseTop && e >= kTopRef) ? t.top + (e - kTopRef) : t.nodes + e;
        const uint4 mt = nd->meta;
        const uint4 lv = nd->live;
        const float4 cb[2] = {nd->box[0], nd->box[1]};
        const uint32_t live[2] = {lv.x, lv.y};
        const uint32_t gamma = lv.z;  // the left range's last leaf
        const uint32_t ref[2] = {mt.x, mt.y};
        const uint32_t first[2] = {mt.z, gamma + 1};
        const uint32_t cnt[2] = {gamma - mt.z + 1, mt.w - gamma};
#pragma unroll
        for (int side = 0; side < 2; side++) {
          if (!node_test<PRED, POINTS>(cb[side], qb)) continue;
          const bool inside = inside_test<PRED, POINTS>(cb[side], qb);
          if (kCounting && inside) {
            inside_live += live[side];
            if (kRecord) {
              const uint32_t f = first[side] | (live[side] == cnt[side] ? kAllMatch : 0u);
              if (nl == 0) {
                lf0 = f;
                ln0 = cnt[side];
              } else {
                lf1 = f;
                ln1 = cnt[side];
              }
              nl++;
            }
          } else if ((ref[side] & kLeafFlag) || cnt[side] <= t.leaf_size || degrade || inside) {
            uint32_t f = first[side];
            if (inside && live[side] == cnt[side]) f |= kAllMatch;
            if (kBound) {
              inside_live += cnt[side];
            } else {
              if (nr == 0) {
                rf0 = f;
                rn0 = cnt[side];
              } else {
                rf1 = f;
                rn1 = cnt[side];
              }
              nr++;
            }
            if (kRecord) {
              if (nl == 0) {
                lf0 = f;
                ln0 = cnt[side];
              } else {
                lf1 = f;
                ln1 = cnt[side];
              }
              nl++;
            }
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
      if (kRecord) {
        const uint32_t linc = wave_inclusive_sum(nl);
        const uint32_t at = nlog + (linc - nl);
        if (nl > 0 && at < log->cap) log_store(log->entries + at, lf0, ln0);
        if (nl > 1 && at + 1 < log->cap) log_store(log->entries + at + 1, lf1, ln1);
        nlog += readlane_u32(linc, kWave - 1);
      }
      lds_fence();
    } else if (LOG && WRITE && m == 0 && lpos < lend) {
      // ---- replay: the next logged ranges into the range list
      const uint32_t c = lend - lpos < (uint32_t)kWaveRanges ? lend - lpos : (uint32_t)kWaveRanges;
      uint32_t carry = 0;  // wave-uniform
      for (uint32_t b0 = 0; b0 < c; b0 += kWave) {
        const uint32_t i = b0 + lane;
        const uint2 en = i < c ? log_load(log->entries + lpos + i) : make_uint2(0u, 0u);
        const uint32_t inc = wave_inclusive_sum(en.y);
        if (i < c) {
          rfirst[i] = en.x;
          rpre[i] = carry + (inc - en.y);
        }
        carry += readlane_u32(inc, kWave - 1);
      }
      if (lane == 0) rpre[c] = carry;
      m = c;
      total = carry;
      lpos += c;
      lds_fence();
    } else if constexpr (kBound) {
      break;  // no ranges are queued: the loop ends with the stack
    } else {
      // ---- scan the queued leaf ranges, kScanStep leaves per step. Lane L
      // takes the leaves at step positions k*64 + L (k < kPerLane), so every
      // load and store instruction of the wave covers 64 consecutive leaves
      // (pairs). The range of each position comes from an LDS map: range
      // starts are marked, a prefix maximum gives each position its range,
      // and the map then holds each position's leaf offset.
      uint32_t e0 = 0;  // wave-uniform: entry holding flat index `base`
      uint16_t* map = L.map;
      for (uint32_t base = 0; base < total; base += kScanStep) {
        const uint32_t lim = total - base < kScanStep ? total - base : kScanStep;
        // (1) clear the step's map, then mark the ranges that start inside
        // the step (entries after e0, in rpre order) with their entry + 1.
#pragma unroll
        for (int k = 0; k < kPerLane; k++) map[k * kWave + lane] = 0;
        lds_fence();
        for (uint32_t e = e0 + 1 + lane;; e += kWave) {
          const bool act = e < m && rpre[e] < base + lim;
          if (act) map[rpre[e] - base] = (uint16_t)(e + 1);
          if (!__any(act)) break;
        }
        lds_fence();
        // (2) owner (entry + 1) of position k*64 + L: the largest mark at or
        // before it, seeded with e0 + 1 -- a max-scan of row k, carried from
        // row to row -- and from it the position's leaf offset.
        uint32_t carry = e0 + 1;  // wave-uniform
        uint32_t e_last = e0;     // wave-uniform: entry of the step's last position
        uint32_t dl[kPerLane];
#pragma unroll
        for (int k = 0; k < kPerLane; k++) {
          uint32_t run = wave_inclusive_max((uint32_t)map[k * kWave + lane]);
          run = run > carry ? run : carry;
          carry = readlane_u32(run, kWave - 1);
          if ((uint32_t)k == (lim - 1) / kWave) e_last = readlane_u32(run, (lim - 1) % kWave) - 1;
          const uint32_t e = run - 1;
          const uint32_t f = rfirst[e];
          dl[k] = ((f & ~kAllMatch) - rpre[e] + kDeltaBias) | (f & kAllMatch);
        }
        e0 = e_last;
        // (3) leaves of positions k*64 + L: ids loaded together with the boxes
        // that need a test, so a step waits for one memory round trip.
        uint32_t leaf[kPerLane];
        uint32_t gid[kPerLane];
        uint32_t valid = 0, allm = 0;
#pragma unroll
        for (int k = 0; k < kPerLane; k++) {
          const uint32_t pos = k * kWave + lane;
          const uint32_t d = dl[k];
          leaf[k] = base + pos + (d & ~kAllMatch) - kDeltaBias;
          const bool ok = pos < lim && leaf[k] < t.n;
          valid |= (ok ? 1u : 0u) << k;
          allm |= ((ok && (d & kAllMatch)) ? 1u : 0u) << k;
          if (WRITE) gid[k] = ok ? t.leaf_ids[leaf[k]] : 0u;
        }
        uint32_t hitmask = allm;
#pragma unroll
        for (int k = 0; k < kPerLane; k++) {
          if (((valid & ~allm) >> k) & 1u)
            hitmask |= (leaf_test<PRED, POINTS>(t, leaf[k], t.leaf_boxes[leaf[k]], qb) ? 1u : 0u)
                       << k;
        }
        // (4) slots: position order is k-major, so the pairs of round k go
        // right after those of rounds < k, in lane order.
        if constexpr (WRITE && HasBeginStep<Emit>::value) {
          uint32_t n = 0;
#pragma unroll
          for (int k = 0; k < kPerLane; k++) n += (uint32_t)__popcll(ballot((hitmask >> k) & 1u));
          if (n) emit.begin_step(emitted, n);
        }
        uint32_t step_hits = 0;  // wave-uniform
#pragma unroll
        for (int k = 0; k < kPerLane; k++) {
          const bool h = (hitmask >> k) & 1u;
          const uint64_t mask = ballot(h);
          if (WRITE && h) emit(emitted + step_hits + mask_rank(mask), gid[k], qid);
          step_hits += (uint32_t)__popcll(mask);
        }
        emitted += step_hits;
        lds_fence();  // the map is rewritten by the next step
      }
      m = 0;
      total = 0;
      lds_fence();
    }
  }
  if (kRecord) log->logged = nlog;
  return emitted + wave_sum_u64(inside_live);
}

// ------------------------------------------------------------- tickets ----
// Persistent waves take queries from one ticket counter per XCD. The query
// order is cut into 8 contiguous slices and workgroups are dealt to the XCDs
// round-robin, so wave w first drains slice w % 8 -- neighbouring (Morton-
// ordered) queries then share one XCD's L2 -- and steals from the other
// slices once its own is empty; costs are heavy-tailed, so no static split
// balances them. The counters take kXcds * kTicketStride words, zeroed
// before the launch.
constexpr uint32_t kXcds = 8;
constexpr uint32_t kTicketStride = 16;  // one 64-byte line per counter

// Next query index for this wave, or >= n when every slice is drained.
// `slice` is the slice to try first; it becomes the one taken from.
__device__ __forceinline__ uint32_t grab_query(uint32_t* tickets, uint32_t n, uint32_t& slice) {
  const uint32_t per = (n + kXcds - 1) / kXcds;
  uint32_t got = 0xFFFFFFFFu, from = slice;
  if (lane_id() == 0) {
    for (uint32_t k = 0; k < kXcds; k++) {
      const uint32_t x = (slice + k) % kXcds;
      const uint32_t begin = x * per;
      const uint32_t end = begin + per < n ? begin + per : n;
      if (begin >= end) continue;
      // No relaxed load ahead of the atomic: across XCDs it is a second
      // memory-side round trip on the same hot line (measured ~1.5x slower).
      const uint32_t v = atomicAdd(tickets + x * kTicketStride, 1u);
      if (v < end - begin) {
        got = begin + v;
        from = x;
        break;
      }
    }
  }
  slice = readlane_u32(from, 0);
  return readlane_u32(got, 0);
}

// ---------------------------------------------------------------- lane ----
// One query per lane with a private stack; for queries with few results
// (points, data-contains-query). Counting mode counts subtrees inside the
// query from their live-leaf counts; writing mode emits the query's pairs in
// traversal order. Returns the query's number of results.
template <int PRED, bool POINTS, bool WRITE, typename Q, typename Emit>
__device__ uint64_t lane_query(const TreeView& t, const Q& qb, const uint32_t qid,
                               const Emit& emit) {
  uint32_t stack[kStackDepth];
  int sp = 0;
  uint32_t cur = 1, end = 0;  // leaf range being scanned: [cur, end]
  bool all = false;           // the range needs no test
  uint64_t found = 0;
  const float4 rb = t.root_box[0];
  if (!node_test<PRED, POINTS>(rb, qb)) return 0;
  const bool root_inside = inside_test<PRED, POINTS>(rb, qb);
  if (root_inside && !WRITE) return t.root_live[0];
  if (t.n <= t.leaf_size || t.n == 1 || root_inside) {
    cur = 0;
    end = t.n - 1;
    all = root_inside && t.root_live[0] == t.n;
  } else {
    stack[sp++] = 0;
  }
  const uint64_t max_steps = 4ull * t.n + 64;
  for (uint64_t steps = 0;; steps++) {
    if (steps > max_steps) {
      atomicOr(t.error, kErrTraversal);
      break;
    }
    if (cur <= end) {
      const uint32_t prim = cur++;
      if (prim < t.n && (all || leaf_test<PRED, POINTS>(t, prim, t.leaf_boxes[prim], qb))) {
        if (WRITE) emit(found, t.leaf_ids[prim], qid);
        found++;
      }
      continue;
    }
    if (sp == 0) break;
    const uint32_t e = stack[--sp];
    if (e & kLeafFlag) {
      cur = end = e & ~kLeafFlag;
      all = false;
      continue;
    }
    if (e & kRangeFlag) {
      const uint32_t node = e & kIndexMask;
      if (node < t.n - 1) {
        const uint4 m = t.nodes[node].meta;
        cur = m.z;
        end = m.w < t.n ? m.w : t.n - 1;
        all = false;
      }
      continue;
    }
    if (e >= t.n - 1) continue;
    const NodeRec* nd = t.nodes + e;
    const uint4 mt = nd->meta;
    const uint4 lv = nd->live;
    const float4 cb[2] = {nd->box[0], nd->box[1]};
    const uint32_t live[2] = {lv.x, lv.y};
    const uint32_t gamma = mt.x & ~kLeafFlag;
    const uint32_t ref[2] = {mt.x, mt.y};
    const uint32_t cnt[2] = {gamma - mt.z + 1, mt.w - gamma};
    // right child first so that the left one is expanded next
#pragma unroll
    for (int k = 0; k < 2; k++) {
      const int side = 1 - k;
      if (!node_test<PRED, POINTS>(cb[side], qb)) continue;
      const bool inside = inside_test<PRED, POINTS>(cb[side], qb);
      if (!WRITE && inside) {
        found += live[side];
      } else if (ref[side] & kLeafFlag) {
        lane_push(stack, sp, ref[side], t.error);
      } else if (cnt[side] <= t.leaf_size || inside) {
        lane_push(stack, sp, ref[side] | kRangeFlag, t.error);
      } else {
        lane_push(stack, sp, ref[side], t.error);
      }
    }
  }
  return found;
}

}  // namespace rts
