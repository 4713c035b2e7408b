// Point<float, 2> (reference include/rtspatial/geom/point.cuh:11-42): two
// floats, 8 bytes, layout-compatible with rts_point; Point<double, 2> below.
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

// Point<double, 2> (reference geom/point.cuh:44-75): two doubles, 16 bytes,
// layout-compatible with rts_point_d. get_coordinate returns the double (the
// reference's returns float, point.cuh:68, so its Envelope predicates on
// doubles compare rounded coordinates; these compare exact ones, as the
// index does by default).
template <>
class Point<double, 2> {
 public:
  using coord_t = double;
  static constexpr int n_dims = 2;
  Point() = default;
  Point(double x, double y) : x_(x), y_(y) {}
  double get_x() const { return x_; }
  double get_y() const { return y_; }
  void set_x(double x) { x_ = x; }
  void set_y(double y) { y_ = y; }
  int get_n_dims() const { return n_dims; }
  double get_coordinate(int dim) const {
    assert(dim < 2);
    return dim == 0 ? x_ : y_;
  }

 private:
  double x_ = 0, y_ = 0;
};

static_assert(sizeof(Point<double, 2>) == 16, "Point<double,2> is two doubles");

}  // namespace rtspatial
#endif
