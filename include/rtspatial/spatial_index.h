// rtspatial::SpatialIndex<float, 2> and <double, 2> -- the reference's C++
// surface (include/rtspatial/spatial_index.cuh:20-931) as a header-only facade
// over the C ABI in rts_c.h. Method names, argument meanings and id semantics
// follow the reference; streams are hipStream_t; failures throw
// rtspatial::Exception (the reference throws on CUDA/OptiX errors and aborts
// on a bad predicate, spatial_index.cuh:387-389).
//
// The result handler of Query(..., void* arg, stream) is chosen at compile
// time, as the reference chooses its shader handler: include
// rtspatial/default_handlers/counting_handlers.h before this header for the
// counting handler (arg: device unsigned long long*), otherwise arg is a
// Queue::DeviceObject() (collecting handler). In a file compiled as HIP,
// QueryWith(pred, queries, handler, stream) takes a handler functor of the
// caller's (rtspatial/device/query.h), and with RTSPATIAL_USER_HANDLERS
// defined, Query(..., void* arg, ...) calls the caller's rtspatial_handle_*
// functions, as a reference build with a custom handler header does.
#ifndef RTSPATIAL_AMD_SPATIAL_INDEX_H
#define RTSPATIAL_AMD_SPATIAL_INDEX_H

#include <hip/hip_runtime_api.h>

#include <cstdint>
#include <iostream>
#include <string>
#include <type_traits>

#include "rts_c.h"
#include "rtspatial/default_handlers/collecting_handlers.h"
#include "rtspatial/geom/envelope.h"
#include "rtspatial/geom/line.h"
#include "rtspatial/geom/point.h"
#include "rtspatial/geom/predicates.h"
#include "rtspatial/utils/array_view.h"
#include "rtspatial/utils/exception.h"
#include "rtspatial/utils/queue.h"
#if defined(__HIP__)
#include "rtspatial/device/query.h"
#endif

namespace rtspatial {

// spatial_index.cuh:22-37. ptx_root, prefer_fast_build_* and compact are
// OptiX settings; they are accepted and ignored.
struct Config {
  std::string ptx_root;
  size_t max_geometries = 1000 * 1000 * 10;
  size_t max_queries = 1000 * 1000;
  bool preallocate = false;
  bool prefer_fast_build_geom = false;
  bool prefer_fast_build_query = false;
  bool compact = false;
  float geom_sample_rate = 0.2;
  float query_sample_rate = 0.2;
  float intersect_cost_weight = 0.95;
  uint32_t max_geom_samples = 10000;
  uint32_t max_query_samples = 1000;
  uint32_t max_parallelism = 512;
  // MI355X additions (see rts_config)
  uint32_t leaf_size = 8;
  uint32_t rebuild_policy = RTS_REBUILD_SAH;
  float sah_rebuild_ratio = 1.5f;
  uint32_t traversal = RTS_TRAVERSAL_AUTO;
  uint32_t sort_queries = 0;
  // SpatialIndex<double, 2> only: RTS_COORD_F64 (exact double predicates) or
  // RTS_COORD_F64_AS_F32 (coordinates rounded to nearest floats: the
  // reference's results, whose double points return float coordinates).
  uint32_t coord = RTS_COORD_F64;
  // Inserts go to a small second tree while they are at most this percentage
  // of the main tree (0: every Insert rebuilds the whole tree).
  uint32_t insert_delta_pct = 2;
};

namespace detail {
// The C entry points of each coordinate type.
template <typename COORD_T>
struct CoordAbi;
template <>
struct CoordAbi<float> {
  using c_point = rts_point;
  using c_envelope = rts_envelope;
  using c_update = rts_update;
  using c_line = rts_line;
  static uint32_t coord(uint32_t) { return RTS_COORD_F32; }
  static int insert(rts_index* i, const c_envelope* b, uint64_t n, hipStream_t s) {
    return rts_index_insert(i, b, n, s);
  }
  static int update(rts_index* i, const c_update* u, uint64_t n, hipStream_t s) {
    return rts_index_update(i, u, n, s);
  }
  static int query(rts_index* i, rts_predicate p, const c_point* q, uint64_t n, rts_handler h,
                   void* a, hipStream_t s) {
    return rts_index_query_points(i, p, q, n, h, a, s);
  }
  static int query(rts_index* i, rts_predicate p, const c_envelope* q, uint64_t n, rts_handler h,
                   void* a, hipStream_t s) {
    return rts_index_query_envelopes(i, p, q, n, h, a, s);
  }
  static int query(rts_index* i, rts_predicate p, const c_line* q, uint64_t n, rts_handler h,
                   void* a, hipStream_t s) {
    return rts_index_query_lines(i, p, q, n, h, a, s);
  }
};
template <>
struct CoordAbi<double> {
  using c_point = rts_point_d;
  using c_envelope = rts_envelope_d;
  using c_update = rts_update_d;
  using c_line = rts_line_d;
  static uint32_t coord(uint32_t requested) { return requested; }
  static int insert(rts_index* i, const c_envelope* b, uint64_t n, hipStream_t s) {
    return rts_index_insert_d(i, b, n, s);
  }
  static int update(rts_index* i, const c_update* u, uint64_t n, hipStream_t s) {
    return rts_index_update_d(i, u, n, s);
  }
  static int query(rts_index* i, rts_predicate p, const c_point* q, uint64_t n, rts_handler h,
                   void* a, hipStream_t s) {
    return rts_index_query_points_d(i, p, q, n, h, a, s);
  }
  static int query(rts_index* i, rts_predicate p, const c_envelope* q, uint64_t n, rts_handler h,
                   void* a, hipStream_t s) {
    return rts_index_query_envelopes_d(i, p, q, n, h, a, s);
  }
  static int query(rts_index* i, rts_predicate p, const c_line* q, uint64_t n, rts_handler h,
                   void* a, hipStream_t s) {
    return rts_index_query_lines_d(i, p, q, n, h, a, s);
  }
};
}  // namespace detail

template <typename COORD_T, int N_DIMS>
class SpatialIndex;

template <typename COORD_T>
class SpatialIndex<COORD_T, 2> {
  using abi = detail::CoordAbi<COORD_T>;

 public:
  using point_t = Point<COORD_T, 2>;
  using envelope_t = Envelope<point_t>;
  using update_t = pair<size_t, envelope_t>;
  using line_t = Line<point_t>;
  static_assert(sizeof(line_t) == sizeof(typename abi::c_line), "line layout");
  static_assert(sizeof(point_t) == sizeof(typename abi::c_point), "point layout");
  static_assert(sizeof(envelope_t) == sizeof(typename abi::c_envelope), "envelope layout");
  static_assert(sizeof(update_t) == sizeof(typename abi::c_update),
                "pair<size_t, envelope> is rts_update / rts_update_d");

  SpatialIndex() = default;
  ~SpatialIndex() {
    if (ix_) rts_index_destroy(ix_);
  }
  SpatialIndex(const SpatialIndex&) = delete;
  SpatialIndex& operator=(const SpatialIndex&) = delete;

  // spatial_index.cuh:83-110
  void Init(const Config& config) {
    if (ix_) {
      rts_index_destroy(ix_);
      ix_ = nullptr;
    }
    rts_config c;
    rts_config_default(&c);
    c.max_geometries = config.max_geometries;
    c.max_queries = config.max_queries;
    c.preallocate = config.preallocate;
    c.prefer_fast_build_geom = config.prefer_fast_build_geom;
    c.prefer_fast_build_query = config.prefer_fast_build_query;
    c.compact = config.compact;
    c.geom_sample_rate = config.geom_sample_rate;
    c.query_sample_rate = config.query_sample_rate;
    c.intersect_cost_weight = config.intersect_cost_weight;
    c.max_geom_samples = config.max_geom_samples;
    c.max_query_samples = config.max_query_samples;
    c.max_parallelism = config.max_parallelism;
    c.leaf_size = config.leaf_size;
    c.rebuild_policy = config.rebuild_policy;
    c.sah_rebuild_ratio = config.sah_rebuild_ratio;
    c.traversal = config.traversal;
    c.sort_queries = config.sort_queries;
    c.coord = abi::coord(config.coord);
    c.insert_delta_pct = config.insert_delta_pct;
    check(rts_index_create(&c, &ix_), "SpatialIndex::Init");
  }

  // spatial_index.cuh:118-129
  void Clear() { check(rts_index_clear(handle()), "SpatialIndex::Clear"); }

  // spatial_index.cuh:142-196
  void Insert(ArrayView<envelope_t> envelopes, hipStream_t stream = nullptr) {
    check(abi::insert(handle(), reinterpret_cast<const typename abi::c_envelope*>(envelopes.data()),
                      envelopes.size(), stream),
          "SpatialIndex::Insert");
  }

  // spatial_index.cuh:208-265
  void Delete(const ArrayView<size_t> ids, hipStream_t stream) {
    static_assert(sizeof(size_t) == sizeof(uint64_t), "size_t ids");
    check(rts_index_delete(handle(), reinterpret_cast<const uint64_t*>(ids.data()), ids.size(),
                           stream),
          "SpatialIndex::Delete");
  }

  // spatial_index.cuh:280-336
  void Update(const ArrayView<update_t> updates, hipStream_t stream) {
    check(abi::update(handle(), reinterpret_cast<const typename abi::c_update*>(updates.data()),
                      updates.size(), stream),
          "SpatialIndex::Update");
  }

  // spatial_index.cuh:352-391
  void Query(Predicate p, ArrayView<point_t> queries, void* arg, hipStream_t stream = nullptr) {
#ifdef RTSPATIAL_USER_HANDLERS
    QueryWithReferenceHandlers(handle(), p, queries.data(), queries.size(), true, kDoubles, arg,
                               stream);
#else
    Query(p, queries, static_cast<rts_handler>(RTSPATIAL_HANDLER), arg, stream);
#endif
  }
  // spatial_index.cuh:407-515
  void Query(Predicate p, ArrayView<envelope_t> queries, void* arg, hipStream_t stream = nullptr) {
#ifdef RTSPATIAL_USER_HANDLERS
    QueryWithReferenceHandlers(handle(), p, queries.data(), queries.size(), false, kDoubles, arg,
                               stream);
#else
    Query(p, queries, static_cast<rts_handler>(RTSPATIAL_HANDLER), arg, stream);
#endif
  }

  // spatial_index.cuh:517-547: kIntersects reports the data boxes a segment
  // meets (exact); kContains reports nothing, as the reference's.
  void Query(Predicate p, ArrayView<line_t> queries, void* arg, hipStream_t stream = nullptr) {
#ifdef RTSPATIAL_USER_HANDLERS
    QueryLinesWithReferenceHandlers(handle(), p, queries.data(), queries.size(), kDoubles, arg,
                                    stream);
#else
    Query(p, queries, static_cast<rts_handler>(RTSPATIAL_HANDLER), arg, stream);
#endif
  }

#if defined(__HIP__)
  // Results to a handler functor of the caller's (rtspatial/device/query.h).
  template <typename Handler>
  void QueryWith(Predicate p, ArrayView<point_t> queries, const Handler& h,
                 hipStream_t stream = nullptr, uint32_t traversal = RTS_TRAVERSAL_AUTO) {
    rtspatial::QueryWith(handle(), p, queries.data(), queries.size(), true, kDoubles, h, stream,
                         traversal);
  }
  template <typename Handler>
  void QueryWith(Predicate p, ArrayView<envelope_t> queries, const Handler& h,
                 hipStream_t stream = nullptr, uint32_t traversal = RTS_TRAVERSAL_AUTO) {
    rtspatial::QueryWith(handle(), p, queries.data(), queries.size(), false, kDoubles, h, stream,
                         traversal);
  }
  template <typename Handler>
  void QueryWith(Predicate p, ArrayView<line_t> queries, const Handler& h,
                 hipStream_t stream = nullptr, uint32_t traversal = RTS_TRAVERSAL_AUTO) {
    rtspatial::QueryLinesWith(handle(), p, queries.data(), queries.size(), kDoubles, h, stream,
                              traversal);
  }
#endif

  // The same with the handler named at run time.
  void Query(Predicate p, ArrayView<point_t> queries, rts_handler handler, void* arg,
             hipStream_t stream) {
    check(abi::query(handle(), static_cast<rts_predicate>(p),
                     reinterpret_cast<const typename abi::c_point*>(queries.data()), queries.size(),
                     handler, arg, stream),
          "SpatialIndex::Query(points)");
  }
  void Query(Predicate p, ArrayView<envelope_t> queries, rts_handler handler, void* arg,
             hipStream_t stream) {
    check(abi::query(handle(), static_cast<rts_predicate>(p),
                     reinterpret_cast<const typename abi::c_envelope*>(queries.data()),
                     queries.size(), handler, arg, stream),
          "SpatialIndex::Query(envelopes)");
  }
  void Query(Predicate p, ArrayView<line_t> queries, rts_handler handler, void* arg,
             hipStream_t stream) {
    check(abi::query(handle(), static_cast<rts_predicate>(p),
                     reinterpret_cast<const typename abi::c_line*>(queries.data()), queries.size(),
                     handler, arg, stream),
          "SpatialIndex::Query(lines)");
  }

  // spatial_index.cuh:766-826 merges the per-Insert BVHs; here it brings the
  // tree up to date and merges the small tree of recent inserts into it.
  void Optimize(hipStream_t stream = nullptr) {
    check(rts_index_optimize(handle(), stream), "Optimize");
  }

  // spatial_index.cuh:893-903
  void PrintMemoryUsage() {
    rts_memory_usage m;
    check(rts_index_memory(handle(), &m), "PrintMemoryUsage");
    const uint64_t mb = 1024 * 1024;
    std::cout << "BVH " << m.bvh / mb << " MB, "
              << "Geometries " << m.geometries / mb << " MB, "
              << "AABBs " << m.aabbs / mb << " MB, "
              << "Total " << m.total / mb << " MB" << std::endl;
  }

  size_t size() const { return ix_ ? rts_index_size(ix_) : 0; }
  void Build(hipStream_t stream = nullptr) { check(rts_index_build(handle(), stream), "Build"); }
  void Sync(hipStream_t stream = nullptr) { check(rts_index_sync(handle(), stream), "Sync"); }
  rts_index* c_handle() { return handle(); }

 private:
  static constexpr bool kDoubles = std::is_same<COORD_T, double>::value;
  rts_index* handle() const {
    if (!ix_) throw Exception("SpatialIndex::Init has not been called");
    return ix_;
  }
  rts_index* ix_ = nullptr;
};

}  // namespace rtspatial
#endif
