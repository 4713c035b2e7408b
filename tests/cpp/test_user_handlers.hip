// Caller-compiled result handlers through the C++ facade, compiled as HIP:
//   * the reference's own handler plugin: rtspatial_handle_* functions
//     defined here (the counting form of default_handlers/
//     counting_handlers.h:7-25) and SpatialIndex::Query(..., void* arg, ...)
//     with RTSPATIAL_USER_HANDLERS, on the reference tests' known answers;
//   * handler functors through SpatialIndex::QueryWith, checked against the
//     library's built-in handlers on the same queries.
#include <hip/hip_runtime.h>

#include <cstdint>

extern "C" __device__ __forceinline__ void rtspatial_handle_point_contains(uint32_t, uint32_t,
                                                                          void* arg) {
  atomicAdd(static_cast<unsigned long long*>(arg), 1ull);
}
extern "C" __device__ __forceinline__ void rtspatial_handle_envelope_contains(uint32_t, uint32_t,
                                                                             void* arg) {
  atomicAdd(static_cast<unsigned long long*>(arg), 1ull);
}
extern "C" __device__ __forceinline__ void rtspatial_handle_envelope_intersects(uint32_t, uint32_t,
                                                                               void* arg) {
  atomicAdd(static_cast<unsigned long long*>(arg), 1ull);
}
#define RTSPATIAL_USER_HANDLERS 1

#include "rtspatial/rtspatial.h"

#include <vector>

#include "mini_test.h"
#include "ref_inputs.h"

using namespace rtspatial;
using ref_inputs::uniform_boxes;
using ref_inputs::uniform_points;
using index_t = SpatialIndex<float, 2>;
using point_t = index_t::point_t;
using envelope_t = index_t::envelope_t;

namespace {

SharedValue<unsigned long long>* g_counter = nullptr;
SharedValue<unsigned long long>& counter() { return *g_counter; }

template <typename Index, typename View>
unsigned long long ref_count(Index& index, Predicate p, View q, Stream& s) {
  counter().set(s.cuda_stream(), 0);
  index.Query(p, q, counter().data(), s.cuda_stream());
  return counter().get(s.cuda_stream());
}

// Per-query counts through done() only, and per-pair counts through
// operator(): both must match the built-in counting handler.
struct PerQueryDone {
  static constexpr bool kPairs = false;
  unsigned long long* out;
  __device__ void done(uint32_t q, uint64_t c) const { out[q] = c; }
};
struct PerQueryPairs {
  unsigned long long* out;
  __device__ void operator()(uint32_t, uint32_t q) const { atomicAdd(out + q, 1ull); }
};

template <typename View>
void check_functors(index_t& index, Predicate p, View q, unsigned long long want, Stream& s) {
  for (uint32_t trav : {(uint32_t)RTS_TRAVERSAL_THREAD, (uint32_t)RTS_TRAVERSAL_WAVE}) {
    device_vector<unsigned long long> a(q.size()), b(q.size());
    check_hip(hipMemsetAsync(a.data(), 0xFF, q.size() * 8, s.cuda_stream()), "memset");
    check_hip(hipMemsetAsync(b.data(), 0, q.size() * 8, s.cuda_stream()), "memset");
    index.QueryWith(p, q, PerQueryDone{a.data()}, s.cuda_stream(), trav);
    index.QueryWith(p, q, PerQueryPairs{b.data()}, s.cuda_stream(), trav);
    s.Sync();
    std::vector<unsigned long long> ha = a.to_host(), hb = b.to_host();
    unsigned long long sa = 0, sb = 0;
    for (size_t i = 0; i < ha.size(); i++) {
      ASSERT_EQ(ha[i], hb[i]);
      sa += ha[i];
      sb += hb[i];
    }
    ASSERT_EQ(sa, want);
    ASSERT_EQ(sb, want);
  }
}

}  // namespace

// test/test_point_queries.h:9-28 with the reference's handler functions.
TEST(UserHandlers, reference_functions_point_known_answer) {
  device_vector<envelope_t> boxes(uniform_boxes(100000, 0.5, 0.5));
  device_vector<point_t> queries(uniform_points(1000));
  Stream s;
  index_t index;
  index.Init(Config());
  index.Insert(ArrayView<envelope_t>(boxes), s.cuda_stream());
  ASSERT_EQ(ref_count(index, Predicate::kContains, ArrayView<point_t>(queries), s), 4224111ull);
  check_functors(index, Predicate::kContains, ArrayView<point_t>(queries), 4224111ull, s);
}

// test/test_envelope_queries.h:38-58 (exact count: 6549770).
TEST(UserHandlers, reference_functions_intersects_known_answer) {
  device_vector<envelope_t> boxes(uniform_boxes(100000, 0.5, 0.5));
  device_vector<envelope_t> queries(uniform_boxes(1000, 0.1, 0.1));
  Stream s;
  index_t index;
  index.Init(Config());
  index.Insert(ArrayView<envelope_t>(boxes), s.cuda_stream());
  ASSERT_EQ(ref_count(index, Predicate::kIntersects, ArrayView<envelope_t>(queries), s),
            6549770ull);
  check_functors(index, Predicate::kIntersects, ArrayView<envelope_t>(queries), 6549770ull, s);
  // contains and within against the built-in counting handler
  for (Predicate p : {Predicate::kContains, Predicate::kWithin}) {
    counter().set(s.cuda_stream(), 0);
    index.Query(p, ArrayView<envelope_t>(queries), RTS_HANDLER_COUNT, counter().data(),
                s.cuda_stream());
    const unsigned long long want = counter().get(s.cuda_stream());
    check_functors(index, p, ArrayView<envelope_t>(queries), want, s);
  }
  counter().set(s.cuda_stream(), 0);
  index.Query(Predicate::kContains, ArrayView<envelope_t>(queries), RTS_HANDLER_COUNT,
              counter().data(), s.cuda_stream());
  const unsigned long long contains = counter().get(s.cuda_stream());
  ASSERT_EQ(ref_count(index, Predicate::kContains, ArrayView<envelope_t>(queries), s), contains);
  bool threw = false;
  try {
    ref_count(index, Predicate::kWithin, ArrayView<envelope_t>(queries), s);
  } catch (const Exception&) {
    threw = true;  // the reference has no within handler function
  }
  ASSERT_TRUE(threw);
}

// test/test_point_queries.h:92-120 on a double index, reference handler functions.
TEST(UserHandlers, reference_functions_fp64_known_answer) {
  using dindex_t = SpatialIndex<double, 2>;
  using dpoint_t = dindex_t::point_t;
  using denv_t = dindex_t::envelope_t;
  pinned_vector<denv_t> boxes;
  pinned_vector<dpoint_t> points;
  boxes.push_back(denv_t(dpoint_t(0, 0), dpoint_t(1, 1)));
  boxes.push_back(denv_t(dpoint_t(0.2, 0.2), dpoint_t(0.8, 0.8)));
  for (auto p : {dpoint_t(0.1, 0.1), dpoint_t(0.5, 0.5), dpoint_t(0.5, 0.6), dpoint_t(1.0, 1.1),
                 dpoint_t(2.0, 2.0)})
    points.push_back(p);
  Stream s;
  dindex_t index;
  index.Init(Config());
  index.Insert(ArrayView<denv_t>(boxes), s.cuda_stream());
  ASSERT_EQ(ref_count(index, Predicate::kContains, ArrayView<dpoint_t>(points), s), 5ull);
}

// Line segments (test/test_line_queries.h:9-29 sizes): the reference's
// handler function and handler functors against the built-in counting
// handler; kContains has no results but done() still reports each query.
TEST(UserHandlers, lines_known_answer) {
  using line_t = index_t::line_t;
  device_vector<envelope_t> boxes(uniform_boxes(100000, 0.5, 0.5));
  device_vector<line_t> lines(ref_inputs::uniform_lines(1000, 0.1, 0.1));
  Stream s;
  index_t index;
  index.Init(Config());
  index.Insert(ArrayView<envelope_t>(boxes), s.cuda_stream());
  ASSERT_EQ(ref_count(index, Predicate::kIntersects, ArrayView<line_t>(lines), s), 6370306ull);
  check_functors(index, Predicate::kIntersects, ArrayView<line_t>(lines), 6370306ull, s);
  check_functors(index, Predicate::kContains, ArrayView<line_t>(lines), 0ull, s);
}

// An empty index: done() still reports every query, with 0 results.
TEST(UserHandlers, empty_index_reports_zero_per_query) {
  device_vector<point_t> queries(uniform_points(100));
  Stream s;
  index_t index;
  index.Init(Config());
  device_vector<unsigned long long> a(100);
  check_hip(hipMemsetAsync(a.data(), 0xFF, 100 * 8, s.cuda_stream()), "memset");
  index.QueryWith(Predicate::kContains, ArrayView<point_t>(queries), PerQueryDone{a.data()},
                  s.cuda_stream());
  s.Sync();
  for (unsigned long long v : a.to_host()) ASSERT_EQ(v, 0ull);
}

int main(int argc, char** argv) {
  int rc;
  {
    SharedValue<unsigned long long> c;
    g_counter = &c;
    rc = mini_test::run_all(argc > 1 ? argv[1] : nullptr);
    g_counter = nullptr;
  }
  return rc;
}
