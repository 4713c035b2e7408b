// Envelope<Point<float,2>> (reference include/rtspatial/geom/envelope.cuh:13-109):
// {min, max}, 16 bytes, layout-compatible with rts_envelope. The predicates
// are the exact closed-interval tests the index evaluates on the GPU.
#ifndef RTSPATIAL_AMD_GEOM_ENVELOPE_H
#define RTSPATIAL_AMD_GEOM_ENVELOPE_H

#include "rtspatial/geom/point.h"

namespace rtspatial {

template <typename POINT_T>
class Envelope {
 public:
  using point_t = POINT_T;
  using coord_t = typename point_t::coord_t;
  static constexpr int n_dims = point_t::n_dims;

  Envelope() = default;
  Envelope(const point_t& min, const point_t& max) : min_(min), max_(max) {}
  const point_t& get_min() const { return min_; }
  const point_t& get_max() const { return max_; }

  // envelope.cuh:29-38
  bool Contains(const point_t& p) const {
    bool c = true;
    for (int d = 0; d < n_dims; d++)
      c &= p.get_coordinate(d) >= min_.get_coordinate(d) && p.get_coordinate(d) <= max_.get_coordinate(d);
    return c;
  }
  // envelope.cuh:40-53
  bool Contains(const Envelope& o) const {
    bool c = true;
    for (int d = 0; d < n_dims; d++)
      c &= o.min_.get_coordinate(d) >= min_.get_coordinate(d) &&
           o.max_.get_coordinate(d) <= max_.get_coordinate(d);
    return c;
  }
  // envelope.cuh:55-69
  bool Intersects(const Envelope& o) const {
    bool c = true;
    for (int d = 0; d < n_dims; d++)
      c &= o.min_.get_coordinate(d) <= max_.get_coordinate(d) &&
           o.max_.get_coordinate(d) >= min_.get_coordinate(d);
    return c;
  }
  // envelope.cuh:103-105
  bool IsValid() const { return min_.get_x() < max_.get_x() && min_.get_y() < max_.get_y(); }

 private:
  point_t min_, max_;
};

static_assert(sizeof(Envelope<Point<float, 2>>) == 16, "Envelope is {min, max}");
static_assert(sizeof(Envelope<Point<double, 2>>) == 32, "Envelope<double> is {min, max}");

}  // namespace rtspatial
#endif
