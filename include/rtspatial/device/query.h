// Queries (points, envelopes, line segments) whose results go to a handler
// compiled by the caller.
//
// The reference compiles its query shaders around a handler header: four
// extern "C" device functions rtspatial_handle_*(geom_id, query_id, arg) that
// its intersection programs call for every result (default_handlers/*.h,
// src/shaders/*.cu, wired in by test/CMakeLists.txt:8). Here the caller's HIP
// code includes this header; QueryWith instantiates the library's traversal
// kernels (rtspatial/device/traverse.h) with the caller's handler and runs
// them on the index's tree (rts_index_tree_view). Compile the including file
// as HIP for gfx950 (hipcc -x hip --offload-arch=gfx950) and link
// librtspatial_amd. Add -mllvm -amdgpu-atomic-optimizer-strategy=DPP when a
// handler does atomics: the lanes of a wave report results of the same query
// together, and that strategy combines their same-address atomics with a
// log-depth reduction instead of a loop over the lanes (a handler with three
// atomics per pair: ~70 ms instead of ~340 ms for 1e10 pairs).
//
// A handler is a trivially copyable functor (it is passed as a kernel
// argument) with
//   __device__ void operator()(uint32_t geom_id, uint32_t query_id) const;
//       called once for every result pair, from the thread that found it;
// and optionally
//   __device__ void done(uint32_t query_id, uint64_t count) const;
//       called once per query, after all of its pairs, with its result count;
//   static constexpr bool kPairs = false;
//       no pairs wanted: only done() is called, and subtrees that lie inside
//       the query are counted from their live-leaf counts without visiting
//       their leaves.
// Calls for different pairs run concurrently and in no fixed order.
//
// With RTSPATIAL_USER_HANDLERS defined, the reference's handler functions
// themselves can be used: define rtspatial_handle_point_contains,
// rtspatial_handle_envelope_contains and rtspatial_handle_envelope_intersects
// (same signatures as the reference) before including this header, and
// SpatialIndex::Query(p, queries, arg, stream) calls them.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <type_traits>
#include <utility>

#include "rts_c.h"
#include "rtspatial/device/traverse.h"
#include "rtspatial/geom/predicates.h"
#include "rtspatial/utils/exception.h"

namespace rts {
namespace user {

template <typename H, typename = void>
struct HasDone : std::false_type {};
template <typename H>
struct HasDone<H, std::void_t<decltype(std::declval<const H&>().done(0u, (uint64_t)0))>>
    : std::true_type {};

template <typename H, typename = void>
struct WantsPairs : std::true_type {};
template <typename H>
struct WantsPairs<H, std::void_t<decltype(H::kPairs)>> : std::integral_constant<bool, H::kPairs> {};

template <typename H>
struct CallEmit {
  H h;
  __device__ void operator()(uint64_t, uint32_t geom_id, uint32_t query_id) const {
    h(geom_id, query_id);
  }
};

constexpr int kLaneThreads = 256;

// One query per wavefront, persistent grid with per-XCD tickets. With
// `order`, the wave that takes ticket k works on query order[k].
template <int PRED, bool POINTS, bool F64, typename H>
__global__ void __launch_bounds__(kWave) k_user_wave(const TreeView t, const void* queries,
                                                     uint32_t n, const uint32_t* order,
                                                     uint32_t* tickets, const H h) {
  __shared__ WaveLds lds;
  uint32_t slice = blockIdx.x % kXcds;
  // Every iteration takes a query or finds none left: n + 1 iterations bound the loop.
  for (uint32_t it = 0; it <= n; it++) {
    const uint32_t k = grab_query(tickets, n, slice);
    if (k >= n) return;
    const uint32_t i = order ? order[k] : k;
    query_t<PRED, F64> q;
    load_q(queries, POINTS, i, q);
    uint64_t c;
    if constexpr (WantsPairs<H>::value)
      c = wave_query<PRED, POINTS, true>(t, q, i, CallEmit<H>{h}, lds);
    else
      c = wave_query<PRED, POINTS, false>(t, q, i, NoEmit{}, lds);
    if constexpr (HasDone<H>::value) {
      if (lane_id() == 0) h.done(i, c);
    }
  }
}

// One query per lane (also the path for an empty tree).
template <int PRED, bool POINTS, bool F64, typename H>
__global__ void __launch_bounds__(kLaneThreads) k_user_lane(const TreeView t, const void* queries,
                                                            uint32_t n, const uint32_t* order,
                                                            const H h) {
  const uint32_t k = blockIdx.x * kLaneThreads + threadIdx.x;
  if (k >= n) return;
  const uint32_t i = order ? order[k] : k;
  uint64_t c = 0;
  if (t.n > 0) {
    query_t<PRED, F64> q;
    load_q(queries, POINTS, i, q);
    if constexpr (WantsPairs<H>::value)
      c = lane_query<PRED, POINTS, true>(t, q, i, CallEmit<H>{h});
    else
      c = lane_query<PRED, POINTS, false>(t, q, i, NoEmit{});
  }
  if constexpr (HasDone<H>::value) h.done(i, c);
}

// (Templates, so that every file that includes this header may instantiate them.)
template <int = 0>
__global__ void k_user_zero(uint32_t* p, uint32_t n) {
  for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0u;
}

// Queries of an index that rounds doubles to floats, rounded to nearest.
template <int = 0>
__global__ void k_user_round(const double* src, uint64_t words, float* dst) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < words;
       i += (uint64_t)gridDim.x * blockDim.x)
    dst[i] = (float)src[i];
}

inline void check_hip(hipError_t e, const char* what) {
  if (e != hipSuccess)
    throw rtspatial::Exception(std::string("QueryWith: ") + what + ": " + hipGetErrorString(e));
}

inline uint32_t persistent_blocks(uint32_t n) {
  int dev = 0, cus = 0;
  check_hip(hipGetDevice(&dev), "hipGetDevice");
  check_hip(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev),
            "hipDeviceGetAttribute");
  uint32_t blocks = (uint32_t)((cus > 0 ? cus : 256) * 32);
  if (n < blocks) blocks = n;
  return (blocks + kXcds - 1) / kXcds * kXcds;
}

template <int PRED, bool POINTS, bool F64, typename H>
void launch(const TreeView& t, const void* q, uint32_t n, const uint32_t* order, bool wave,
            const H& h, hipStream_t stream) {
  if (wave && t.n > 0) {
    uint32_t* tickets = nullptr;
    check_hip(hipMallocAsync(reinterpret_cast<void**>(&tickets),
                             kXcds * kTicketStride * sizeof(uint32_t), stream),
              "hipMallocAsync");
    hipLaunchKernelGGL(k_user_zero<>, dim3(1), dim3(128), 0, stream, tickets, kXcds * kTicketStride);
    hipLaunchKernelGGL((k_user_wave<PRED, POINTS, F64, H>), dim3(persistent_blocks(n)),
                       dim3(kWave), 0, stream, t, q, n, order, tickets, h);
    const hipError_t e = hipGetLastError();
    check_hip(hipFreeAsync(tickets, stream), "hipFreeAsync");
    check_hip(e, "launch");
  } else {
    hipLaunchKernelGGL((k_user_lane<PRED, POINTS, F64, H>),
                       dim3((n + kLaneThreads - 1) / kLaneThreads), dim3(kLaneThreads), 0, stream,
                       t, q, n, order, h);
    check_hip(hipGetLastError(), "launch");
  }
}

// kind: an rts_query_kind. Segment queries with kContains have no results;
// they run on an empty tree so that done() still reports each query.
template <bool F64, typename H>
void dispatch(const TreeView& t, int pred, int kind, const void* q, uint32_t n,
              const uint32_t* order, bool wave, const H& h, hipStream_t stream) {
  if (kind == RTS_QUERY_POINTS)
    return launch<RTS_CONTAINS, true, F64>(t, q, n, order, wave, h, stream);
  if (kind == RTS_QUERY_LINES) {
    TreeView tt = t;
    if (pred != RTS_INTERSECTS) tt.n = 0;
    return launch<kPredLine, false, F64>(tt, q, n, order, wave, h, stream);
  }
  switch (pred) {
    case RTS_INTERSECTS:
      return launch<RTS_INTERSECTS, false, F64>(t, q, n, order, wave, h, stream);
    case RTS_CONTAINS: return launch<RTS_CONTAINS, false, F64>(t, q, n, order, wave, h, stream);
    default: return launch<RTS_WITHIN, false, F64>(t, q, n, order, wave, h, stream);
  }
}

}  // namespace user
}  // namespace rts

namespace rtspatial {

namespace detail {

// QueryWith over one query kind (an rts_query_kind).
template <typename Handler>
void QueryWithKind(rts_index* index, int pred, const void* queries, uint64_t n, int kind,
                   bool doubles, const Handler& h, hipStream_t stream, uint32_t traversal) {
  static_assert(std::is_trivially_copyable<Handler>::value, "a handler is a kernel argument");
  if (pred < RTS_INTERSECTS || pred > RTS_WITHIN) throw Exception("QueryWith: unknown predicate");
  if (kind == RTS_QUERY_POINTS && pred != RTS_CONTAINS)
    throw Exception("QueryWith: point queries support only kContains (spatial_index.cuh:385-389)");
  if (kind == RTS_QUERY_LINES && pred == RTS_WITHIN)
    throw Exception("QueryWith: segment queries take kIntersects or kContains "
                    "(spatial_index.cuh:523-545)");
  if (n >= (1ull << 32)) throw Exception("QueryWith: at most 2^32 - 1 queries per call");
  if (n && !queries) throw Exception("QueryWith: null queries");
  rts_tree_view v;
  check(rts_index_tree_view(index, stream, &v), "QueryWith");
  rts_query_plan plan;
  check(rts_index_query_plan(index, static_cast<rts_predicate>(pred), kind, n, &plan),
        "QueryWith");
  if (doubles != (v.coord != RTS_COORD_F32))
    throw Exception(doubles ? "QueryWith: double queries on a float index"
                            : "QueryWith: float queries on a double index");
  if (n == 0) return;
  rts::TreeView t;
  t.leaf_boxes = static_cast<const float4*>(v.leaf_boxes);
  t.leaf_ids = v.leaf_ids;
  t.nodes = static_cast<const rts::NodeRec*>(v.nodes);
  t.root_box = static_cast<const float4*>(v.root_box);
  t.root_live = v.root_live;
  t.n = v.n;
  t.leaf_size = plan.leaf_size;
  t.error = v.error;
  t.boxes64 = v.coord == RTS_COORD_F64 ? static_cast<const rts::Box64*>(v.boxes64) : nullptr;
  const bool wave = traversal == RTS_TRAVERSAL_AUTO ? plan.traversal == RTS_TRAVERSAL_WAVE
                                                    : traversal == RTS_TRAVERSAL_WAVE;
  uint32_t* order = nullptr;
  if (plan.sort_queries == 1 && v.n > 0) {
    rts::user::check_hip(hipMallocAsync(reinterpret_cast<void**>(&order), n * sizeof(uint32_t),
                                        stream),
                         "hipMallocAsync");
    const int rc = rts_index_order_queries(index, queries, n, kind, order, stream);
    if (rc != RTS_OK) {
      (void)hipFreeAsync(order, stream);
      check(rc, "QueryWith");
    }
  }
  if (v.coord == RTS_COORD_F64) {
    rts::user::dispatch<true>(t, pred, kind, queries, (uint32_t)n, order, wave, h, stream);
  } else if (v.coord == RTS_COORD_F64_AS_F32) {
    const uint64_t words = n * (kind == RTS_QUERY_POINTS ? 2 : 4);
    float* rounded = nullptr;
    rts::user::check_hip(hipMallocAsync(reinterpret_cast<void**>(&rounded), words * sizeof(float),
                                        stream),
                         "hipMallocAsync");
    uint64_t blocks = (words + 255) / 256;
    hipLaunchKernelGGL(rts::user::k_user_round<>, dim3((uint32_t)(blocks < 4096 ? blocks : 4096)),
                       dim3(256), 0, stream, static_cast<const double*>(queries), words, rounded);
    rts::user::dispatch<false>(t, pred, kind, rounded, (uint32_t)n, order, wave, h, stream);
    rts::user::check_hip(hipFreeAsync(rounded, stream), "hipFreeAsync");
  } else {
    rts::user::dispatch<false>(t, pred, kind, queries, (uint32_t)n, order, wave, h, stream);
  }
  if (order) rts::user::check_hip(hipFreeAsync(order, stream), "hipFreeAsync");
}

}  // namespace detail

// Runs `queries` (n points or envelopes, floats for a float index, doubles
// for a double one, in device or pinned memory) against `index` with handler
// `h` on `stream`, under the plan the library would run the same call under
// (rts_index_query_plan: traversal mode, Morton ordering, leaf-range size);
// `traversal` other than RTS_TRAVERSAL_AUTO overrides its mode. Throws
// rtspatial::Exception on a bad argument or a failed launch; rts_index_sync
// reports a kernel's internal error.
template <typename Handler>
void QueryWith(rts_index* index, Predicate p, const void* queries, uint64_t n, bool points,
               bool doubles, const Handler& h, hipStream_t stream = nullptr,
               uint32_t traversal = RTS_TRAVERSAL_AUTO) {
  detail::QueryWithKind(index, static_cast<int>(p), queries, n,
                        points ? RTS_QUERY_POINTS : RTS_QUERY_ENVELOPES, doubles, h, stream,
                        traversal);
}

// The same for n line segments (rts_line / rts_line_d layout): kIntersects
// calls the handler for every data box a segment meets; kContains has no
// results (spatial_index.cuh:524-526), so only done() is called, with 0.
template <typename Handler>
void QueryLinesWith(rts_index* index, Predicate p, const void* queries, uint64_t n, bool doubles,
                    const Handler& h, hipStream_t stream = nullptr,
                    uint32_t traversal = RTS_TRAVERSAL_AUTO) {
  detail::QueryWithKind(index, static_cast<int>(p), queries, n, RTS_QUERY_LINES, doubles, h,
                        stream, traversal);
}

#ifdef RTSPATIAL_USER_HANDLERS
// The reference's handler functions as handlers (default_handlers/*.h): the
// caller defines them before including this header.
struct RefPointContains {
  void* arg;
  __device__ void operator()(uint32_t g, uint32_t q) const {
    rtspatial_handle_point_contains(g, q, arg);
  }
};
struct RefEnvelopeContains {
  void* arg;
  __device__ void operator()(uint32_t g, uint32_t q) const {
    rtspatial_handle_envelope_contains(g, q, arg);
  }
};
struct RefEnvelopeIntersects {
  void* arg;
  __device__ void operator()(uint32_t g, uint32_t q) const {
    rtspatial_handle_envelope_intersects(g, q, arg);
  }
};

// spatial_index.cuh:352-515 with the caller's handler functions. The
// reference has no handler for kWithin; it is rejected here.
inline void QueryWithReferenceHandlers(rts_index* index, Predicate p, const void* queries,
                                       uint64_t n, bool points, bool doubles, void* arg,
                                       hipStream_t stream) {
  if (points)
    return QueryWith(index, p, queries, n, true, doubles, RefPointContains{arg}, stream);
  if (p == Predicate::kContains)
    return QueryWith(index, p, queries, n, false, doubles, RefEnvelopeContains{arg}, stream);
  if (p == Predicate::kIntersects)
    return QueryWith(index, p, queries, n, false, doubles, RefEnvelopeIntersects{arg}, stream);
  throw Exception("Query: the reference's handler functions have no kWithin form");
}

// spatial_index.cuh:517-547 with the caller's handler functions: segment
// intersects results go to rtspatial_handle_envelope_intersects, the function
// the reference's line program is written to call
// (shaders_intersects_line_query_2d.cu:36); kContains has no results.
inline void QueryLinesWithReferenceHandlers(rts_index* index, Predicate p, const void* queries,
                                            uint64_t n, bool doubles, void* arg,
                                            hipStream_t stream) {
  return QueryLinesWith(index, p, queries, n, doubles, RefEnvelopeIntersects{arg}, stream);
}
#endif

}  // namespace rtspatial
