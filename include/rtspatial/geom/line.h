// Line<Point<COORD_T,2>> (reference include/rtspatial/geom/line.cuh:10-31):
// the segment p1-p2, {p1, p2}, layout-compatible with rts_line (float) and
// rts_line_d (double).
#ifndef RTSPATIAL_AMD_GEOM_LINE_H
#define RTSPATIAL_AMD_GEOM_LINE_H

#include "rtspatial/geom/point.h"

namespace rtspatial {

template <typename POINT_T>
class Line {
 public:
  using point_t = POINT_T;
  using coord_t = typename point_t::coord_t;

  Line() = default;
  Line(const point_t& p1, const point_t& p2) : p1_(p1), p2_(p2) {}
  const point_t& get_p1() const { return p1_; }
  const point_t& get_p2() const { return p2_; }

 private:
  point_t p1_, p2_;
};

static_assert(sizeof(Line<Point<float, 2>>) == 16, "Line is {p1, p2}");
static_assert(sizeof(Line<Point<double, 2>>) == 32, "Line<double> is {p1, p2}");

}  // namespace rtspatial
#endif
