// Point<float, 2> (reference include/rtspatial/geom/point.cuh:11-42): two
// floats, 8 bytes, layout-compatible with rts_point.
#ifndef RTSPATIAL_AMD_GEOM_POINT_H
#define RTSPATIAL_AMD_GEOM_POINT_H

#include <cassert>

namespace rtspatial {

template <typename COORD_T, int N_DIMS>
class Point;

template <>
class Point<float, 2> {
 public:
  using coord_t = float;
  static constexpr int n_dims = 2;
  Point() = default;
  Point(float x, float y) : x_(x), y_(y) {}
  float get_x() const { return x_; }
  float get_y() const { return y_; }
  void set_x(float x) { x_ = x; }
  void set_y(float y) { y_ = y; }
  int get_n_dims() const { return n_dims; }
  float get_coordinate(int dim) const {
    assert(dim < 2);
    return dim == 0 ? x_ : y_;
  }

 private:
  float x_ = 0, y_ = 0;
};

static_assert(sizeof(Point<float, 2>) == 8, "Point<float,2> is two floats");

}  // namespace rtspatial
#endif
