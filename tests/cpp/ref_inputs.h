// Input generators of the reference's tests, shared by the C++ tests.
#pragma once

#include <algorithm>
#include <random>
#include <vector>

#include "rtspatial/geom/envelope.h"
#include "rtspatial/geom/line.h"
#include "rtspatial/geom/point.h"

namespace ref_inputs {

using point_t = rtspatial::Point<float, 2>;
using envelope_t = rtspatial::Envelope<point_t>;

// The reference tests' input streams (test/test_commons.h:22-67): std::mt19937
// seeded with the element count, libstdc++ uniform_real_distribution<float>.
inline std::vector<envelope_t> uniform_boxes(size_t n, float max_w, float max_h) {
  std::mt19937 mt(n);
  std::uniform_real_distribution<float> min_x(0.0, 1.0), min_y(0.0, 1.0);
  std::uniform_real_distribution<float> w_dist(0.0, max_w), h_dist(0.0, max_h);
  std::vector<envelope_t> out(n);
  for (size_t i = 0; i < n; i++) {
    float x = min_x(mt), y = min_y(mt), w = w_dist(mt), h = h_dist(mt);
    out[i] = envelope_t(point_t(x, y), point_t(std::min(x + w, 1.0f), std::min(y + h, 1.0f)));
  }
  return out;
}

inline std::vector<point_t> uniform_points(size_t n) {
  std::mt19937 mt(n);
  std::uniform_real_distribution<float> gx(0.0, 1.0), gy(0.0, 1.0);
  std::vector<point_t> out(n);
  for (size_t i = 0; i < n; i++) {
    float x = gx(mt), y = gy(mt);
    out[i] = point_t(x, y);
  }
  return out;
}

// GenerateUniformLines (test/test_commons.h:69-97): the same stream as
// uniform_boxes, each segment the diagonal from (x, y) to the clipped corner.
inline std::vector<rtspatial::Line<point_t>> uniform_lines(size_t n, float max_w, float max_h) {
  std::vector<rtspatial::Line<point_t>> out;
  for (const envelope_t& e : uniform_boxes(n, max_w, max_h))
    out.push_back(rtspatial::Line<point_t>(e.get_min(), e.get_max()));
  return out;
}

}  // namespace ref_inputs
