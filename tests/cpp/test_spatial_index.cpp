// The reference's test suite (test/test_point_queries.h,
// test/test_envelope_queries.h) against the C++ facade, built with the
// counting handler as the reference's tests are (test/CMakeLists.txt:8-9).
// Counts are the exact closed-predicate answers: the reference's 6549771 for
// the 100k x 1k intersects case includes one pair its ray formulation admits
// with a 1-ulp gap (query 362, box 76881); the exact count is 6549770.
#include "rtspatial/default_handlers/counting_handlers.h"
#include "rtspatial/rtspatial.h"

#include <algorithm>
#include <random>
#include <unordered_set>

#include "mini_test.h"
#include "ref_inputs.h"

using namespace rtspatial;
using ref_inputs::uniform_boxes;
using ref_inputs::uniform_points;
using index_t = SpatialIndex<float, 2>;
using point_t = index_t::point_t;
using envelope_t = index_t::envelope_t;

namespace {

// The counting handler's device counter; created in main() and released
// before main() returns (no device memory is freed during static teardown).
SharedValue<unsigned long long>* g_counter = nullptr;
SharedValue<unsigned long long>& counter() { return *g_counter; }

template <typename Q>
unsigned long long count(index_t& index, Predicate p, device_vector<Q>& q, Stream& s) {
  counter().set(s.cuda_stream(), 0);
  index.Query(p, ArrayView<Q>(q), counter().data(), s.cuda_stream());
  return counter().get(s.cuda_stream());
}

void insert_batches(index_t& index, device_vector<envelope_t>& boxes, int n_batches, Stream& s) {
  size_t avg = (boxes.size() + n_batches - 1) / n_batches;
  for (int b = 0; b < n_batches; b++) {
    size_t begin = b * avg, end = std::min(begin + avg, boxes.size());
    index.Insert(ArrayView<envelope_t>(boxes.data() + begin, end - begin), s.cuda_stream());
  }
}

}  // namespace

// test/test_point_queries.h:9-28
TEST(PointQueries, fp32_contains_point_large) {
  device_vector<envelope_t> boxes(uniform_boxes(100000, 0.5, 0.5));
  device_vector<point_t> queries(uniform_points(1000));
  Stream s;
  index_t index;
  index.Init(Config());
  index.Insert(ArrayView<envelope_t>(boxes), s.cuda_stream());
  ASSERT_EQ(count(index, Predicate::kContains, queries, s), 4224111ull);
}

// test/test_point_queries.h:30-60
TEST(PointQueries, fp32_contains_point_large_batch) {
  device_vector<envelope_t> boxes(uniform_boxes(100000, 0.5, 0.5));
  device_vector<point_t> queries(uniform_points(1000));
  Stream s;
  index_t index;
  index.Init(Config());
  insert_batches(index, boxes, 10, s);
  ASSERT_EQ(count(index, Predicate::kContains, queries, s), 4224111ull);
}

// test/test_point_queries.h:62-90: pinned host inputs
TEST(PointQueries, fp32_contains_point) {
  pinned_vector<envelope_t> boxes;
  pinned_vector<point_t> points;
  boxes.push_back(envelope_t(point_t(0, 0), point_t(1, 1)));
  boxes.push_back(envelope_t(point_t(0.2, 0.2), point_t(0.8, 0.8)));
  for (auto p : {point_t(0.1, 0.1), point_t(0.5, 0.5), point_t(0.5, 0.6), point_t(1.0, 1.1),
                 point_t(2.0, 2.0)})
    points.push_back(p);
  Stream s;
  index_t index;
  index.Init(Config());
  index.Insert(ArrayView<envelope_t>(boxes), s.cuda_stream());
  counter().set(s.cuda_stream(), 0);
  index.Query(Predicate::kContains, ArrayView<point_t>(points), counter().data(), s.cuda_stream());
  ASSERT_EQ(counter().get(s.cuda_stream()), 5ull);
}

// test/test_envelope_queries.h:11-36
TEST(EnvelopeQueries, fp32_intersects_envelope) {
  pinned_vector<envelope_t> boxes, queries;
  boxes.push_back(envelope_t(point_t(0, 0), point_t(1, 1)));
  boxes.push_back(envelope_t(point_t(0.5, -0.5), point_t(1.5, 0.5)));
  queries.push_back(envelope_t(point_t(0.75, 0.25), point_t(1.25, 0.75)));
  Stream s;
  index_t index;
  index.Init(Config());
  index.Insert(ArrayView<envelope_t>(boxes), s.cuda_stream());
  counter().set(s.cuda_stream(), 0);
  index.Query(Predicate::kIntersects, ArrayView<envelope_t>(queries), counter().data(),
              s.cuda_stream());
  ASSERT_EQ(counter().get(s.cuda_stream()), 2ull);
}

// test/test_envelope_queries.h:38-58 (exact answer, see the file comment)
TEST(EnvelopeQueries, fp32_intersects_envelope_big) {
  device_vector<envelope_t> boxes(uniform_boxes(100000, 0.5, 0.5));
  device_vector<envelope_t> queries(uniform_boxes(1000, 0.1, 0.1));
  Stream s;
  index_t index;
  index.Init(Config());
  index.Insert(ArrayView<envelope_t>(boxes), s.cuda_stream());
  ASSERT_EQ(count(index, Predicate::kIntersects, queries, s), 6549770ull);
}

// test/test_envelope_queries.h:60-91
TEST(EnvelopeQueries, fp32_intersects_envelope_batch) {
  device_vector<envelope_t> boxes(uniform_boxes(100000, 0.5, 0.5));
  device_vector<envelope_t> queries(uniform_boxes(1000, 0.1, 0.1));
  Stream s;
  index_t index;
  index.Init(Config());
  insert_batches(index, boxes, 10, s);
  ASSERT_EQ(count(index, Predicate::kIntersects, queries, s), 6549770ull);
  // Optimize (spatial_index.cuh:766-826) changes no result; PrintMemoryUsage
  // (:893-903) reports the index's memory.
  index.Optimize(s.cuda_stream());
  ASSERT_EQ(count(index, Predicate::kIntersects, queries, s), 6549770ull);
  index.PrintMemoryUsage();
}

// test/test_envelope_queries.h:93-164: refit after Update == a fresh build
TEST(EnvelopeQueries, fp32_intersects_envelope_batch_update) {
  const size_t n1 = 100000;
  auto host_boxes = uniform_boxes(n1, 0.5, 0.5);
  device_vector<envelope_t> boxes(host_boxes);
  device_vector<envelope_t> queries(uniform_boxes(1000, 0.1, 0.1));
  Stream s;
  index_t index;
  index.Init(Config());
  insert_batches(index, boxes, 10, s);
  const size_t n_updates = n1 * 0.05;
  auto new_boxes = uniform_boxes(n_updates, 0.5, 0.5);
  std::vector<index_t::update_t> updates;
  std::unordered_set<size_t> ids;
  std::uniform_int_distribution<size_t> id_dist(0, n1 - 1);
  std::mt19937 mt(n_updates);
  while (ids.size() < n_updates) {
    size_t id = id_dist(mt);
    if (ids.find(id) == ids.end()) {
      updates.push_back({id, new_boxes[ids.size()]});
      ids.insert(id);
    }
  }
  unsigned long long answer;
  {
    for (auto& u : updates) host_boxes[u.first] = u.second;
    device_vector<envelope_t> tmp(host_boxes);
    index_t fresh;
    fresh.Init(Config());
    fresh.Insert(ArrayView<envelope_t>(tmp), s.cuda_stream());
    answer = count(fresh, Predicate::kIntersects, queries, s);
  }
  device_vector<index_t::update_t> d_updates(updates);
  index.Update(ArrayView<index_t::update_t>(d_updates), s.cuda_stream());
  s.Sync();
  ASSERT_EQ(count(index, Predicate::kIntersects, queries, s), answer);
  ASSERT_EQ(answer, 6553263ull);  // the oracle's count on the updated array
}

// test/test_envelope_queries.h:166-245 (with and without `compact`)
TEST(EnvelopeQueries, fp32_test_delete) {
  for (bool compact : {false, true}) {
    pinned_vector<envelope_t> boxes, queries;
    boxes.push_back(envelope_t(point_t(0, 0), point_t(1, 1)));
    boxes.push_back(envelope_t(point_t(0.5, -0.5), point_t(1.5, 0.5)));
    queries.push_back(envelope_t(point_t(0.75, 0.25), point_t(1.25, 0.75)));
    Stream s;
    Config config;
    config.compact = compact;
    index_t index;
    index.Init(config);
    index.Insert(ArrayView<envelope_t>(boxes), s.cuda_stream());
    counter().set(s.cuda_stream(), 0);
    index.Query(Predicate::kIntersects, ArrayView<envelope_t>(queries), counter().data(),
                s.cuda_stream());
    ASSERT_EQ(counter().get(s.cuda_stream()), 2ull);
    device_vector<size_t> deleted(std::vector<size_t>{1});
    index.Delete(ArrayView<size_t>(deleted), s.cuda_stream());
    counter().set(s.cuda_stream(), 0);
    s.Sync();
    index.Query(Predicate::kIntersects, ArrayView<envelope_t>(queries), counter().data(),
                s.cuda_stream());
    ASSERT_EQ(counter().get(s.cuda_stream()), 1ull);
  }
}

// The collecting handler through the run-time overload: the pairs themselves.
TEST(EnvelopeQueries, fp32_collect_pairs) {
  pinned_vector<envelope_t> boxes, queries;
  boxes.push_back(envelope_t(point_t(0, 0), point_t(1, 1)));
  boxes.push_back(envelope_t(point_t(0.5, -0.5), point_t(1.5, 0.5)));
  boxes.push_back(envelope_t(point_t(3, 3), point_t(4, 4)));
  queries.push_back(envelope_t(point_t(0.75, 0.25), point_t(1.25, 0.75)));
  queries.push_back(envelope_t(point_t(3.5, 3.5), point_t(5, 5)));
  Stream s;
  index_t index;
  index.Init(Config());
  index.Insert(ArrayView<envelope_t>(boxes), s.cuda_stream());
  Queue q;
  q.Init(16);
  index.Query(Predicate::kIntersects, ArrayView<envelope_t>(queries), RTS_HANDLER_COLLECT,
              q.DeviceObject(), s.cuda_stream());
  auto got = q.ToHost(s.cuda_stream());
  std::sort(got.begin(), got.end(), [](const Queue::value_type& a, const Queue::value_type& b) {
    return a.second != b.second ? a.second < b.second : a.first < b.first;
  });
  ASSERT_EQ(got.size(), (size_t)3);
  ASSERT_EQ(got[0].first, 0u);
  ASSERT_EQ(got[0].second, 0u);
  ASSERT_EQ(got[1].first, 1u);
  ASSERT_EQ(got[1].second, 0u);
  ASSERT_EQ(got[2].first, 2u);
  ASSERT_EQ(got[2].second, 1u);
}

// test/test_point_queries.h:92-120: the same case on a double index.
TEST(PointQueries, fp64_contains_point) {
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
  counter().set(s.cuda_stream(), 0);
  index.Query(Predicate::kContains, ArrayView<dpoint_t>(points), counter().data(), s.cuda_stream());
  ASSERT_EQ(counter().get(s.cuda_stream()), 5ull);
}

// A point 1e-12 right of a box edge: outside in double, inside once both are
// rounded to float -- the answer of the reference's double index, whose
// points return float coordinates (geom/point.cuh:68).
TEST(PointQueries, fp64_exact_and_rounded) {
  using dindex_t = SpatialIndex<double, 2>;
  using dpoint_t = dindex_t::point_t;
  using denv_t = dindex_t::envelope_t;
  pinned_vector<denv_t> boxes;
  pinned_vector<dpoint_t> points;
  boxes.push_back(denv_t(dpoint_t(0.0, 0.0), dpoint_t(0.3, 0.3)));
  points.push_back(dpoint_t(0.3 + 1e-12, 0.1));
  points.push_back(dpoint_t(0.3, 0.1));
  Stream s;
  for (uint32_t coord : {(uint32_t)RTS_COORD_F64, (uint32_t)RTS_COORD_F64_AS_F32}) {
    dindex_t index;
    Config c;
    c.coord = coord;
    index.Init(c);
    index.Insert(ArrayView<denv_t>(boxes), s.cuda_stream());
    counter().set(s.cuda_stream(), 0);
    index.Query(Predicate::kContains, ArrayView<dpoint_t>(points), counter().data(),
                s.cuda_stream());
    ASSERT_EQ(counter().get(s.cuda_stream()), coord == RTS_COORD_F64 ? 1ull : 2ull);
  }
  ASSERT_TRUE(!boxes[0].Contains(points[0]) && boxes[0].Contains(points[1]));
}

TEST(PointQueries, bad_predicate_throws) {
  pinned_vector<envelope_t> boxes;
  boxes.push_back(envelope_t(point_t(0, 0), point_t(1, 1)));
  pinned_vector<point_t> pts;
  pts.push_back(point_t(0.5, 0.5));
  index_t index;
  index.Init(Config());
  index.Insert(ArrayView<envelope_t>(boxes));
  bool threw = false;
  try {
    index.Query(Predicate::kIntersects, ArrayView<point_t>(pts), counter().data());
  } catch (const Exception&) {
    threw = true;
  }
  ASSERT_TRUE(threw);
}

// test/test_line_queries.h:9-29 (100K boxes, 1000 lines; the reference's
// call omits the line generator's sizes, taken here as its envelope tests'
// 0.1 x 0.1). The reference's line program reports nothing; 6370306 is the
// exact count of box-segment contacts, from the oracle's exact join.
TEST(LineQueries, fp32_intersects) {
  using line_t = index_t::line_t;
  device_vector<envelope_t> boxes(uniform_boxes(100000, 0.5, 0.5));
  device_vector<line_t> lines(ref_inputs::uniform_lines(1000, 0.1, 0.1));
  Stream s;
  index_t index;
  index.Init(Config());
  index.Insert(ArrayView<envelope_t>(boxes), s.cuda_stream());
  ASSERT_EQ(count(index, Predicate::kIntersects, lines, s), 6370306ull);
  ASSERT_EQ(count(index, Predicate::kContains, lines, s), 0ull);  // spatial_index.cuh:524-526
  bool threw = false;
  try {
    count(index, Predicate::kWithin, lines, s);
  } catch (const Exception&) {
    threw = true;
  }
  ASSERT_TRUE(threw);
}

// Closed-form cases on a double index: the unit box against segments that
// cross it, touch a corner, run along an edge, or pass beside it.
TEST(LineQueries, fp64_closed_form) {
  using dindex_t = SpatialIndex<double, 2>;
  using dpoint_t = dindex_t::point_t;
  using denv_t = dindex_t::envelope_t;
  using dline_t = dindex_t::line_t;
  pinned_vector<denv_t> boxes;
  boxes.push_back(denv_t(dpoint_t(0, 0), dpoint_t(1, 1)));
  pinned_vector<dline_t> lines;
  lines.push_back(dline_t(dpoint_t(-1, -1), dpoint_t(2, 2)));        // crosses
  lines.push_back(dline_t(dpoint_t(0, 2), dpoint_t(2, 0)));          // touches (1, 1)
  lines.push_back(dline_t(dpoint_t(1, -3), dpoint_t(1, 3)));         // along the right edge
  lines.push_back(dline_t(dpoint_t(1, 2), dpoint_t(2, 1)));          // beside the corner
  lines.push_back(dline_t(dpoint_t(0, 2.0000000001), dpoint_t(2.0000000001, 0)));  // 1e-10 off
  lines.push_back(dline_t(dpoint_t(0.5, 0.5), dpoint_t(0.5, 0.5)));  // a point inside
  Stream s;
  dindex_t index;
  index.Init(Config());
  index.Insert(ArrayView<denv_t>(boxes), s.cuda_stream());
  counter().set(s.cuda_stream(), 0);
  index.Query(Predicate::kIntersects, ArrayView<dline_t>(lines), counter().data(),
              s.cuda_stream());
  ASSERT_EQ(counter().get(s.cuda_stream()), 4ull);
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
