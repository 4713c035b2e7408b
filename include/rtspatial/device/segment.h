// Segment-box intersection for line queries, exact.
//
// A closed segment p1-p2 meets a closed box iff (separating axes of a segment
// and a box) their bounding boxes overlap and the box's four corners do not
// all lie strictly on one side of the line through p1 and p2. The side of a
// corner c is the sign of the orientation
//   orient(p1, p2, c) = p1x p2y - p1y p2x + p2x cy - p2y cx + cx p1y - cy p1x,
// evaluated exactly: each product is split into its rounded value and its
// error (a two-product through fma), and the twelve doubles are summed into a
// nonoverlapping expansion whose largest component carries the sign
// (Shewchuk, "Adaptive Precision Floating-Point Arithmetic and Fast Robust
// Geometric Predicates", 1997: Two-Sum, Two-Product, Grow-Expansion). A
// segment whose ends coincide has orientation 0 everywhere and reduces to the
// bounding-box test. Exact for any float coordinates (a product of two
// floats is exact in double) and for double coordinates that are 0 or of
// magnitude within 2^-450 .. 2^500 (products neither overflow nor round an
// error term below the normal range).
//
// Only two corners need a side: those where the orientation is largest and
// smallest. Each is first evaluated in double with a bound on its rounding
// error; the exact sum runs only when the double value is within the bound.
// Node culling uses the double evaluation alone, as a necessary condition (a
// corner counts as strictly on a side only when its value exceeds the bound).
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

namespace rts {

// Every function below evaluates exactly the operations written (no fused
// multiply-adds formed by the compiler): the error terms depend on it.
__host__ __device__ __forceinline__ void two_sum(double a, double b, double& x, double& y) {
#pragma clang fp contract(off)
  x = a + b;
  const double bv = x - a;
  const double av = x - bv;
  y = (a - av) + (b - bv);
}

__host__ __device__ __forceinline__ void two_prod(double a, double b, double& x, double& y) {
#pragma clang fp contract(off)
  x = a * b;
  y = fma(a, b, -x);
}

// Sign (-1, 0, 1) of the exact orientation of c against the line p1 -> p2.
// Out of line: it runs only when the double evaluation is within its error
// bound, and inlined it would raise the register count of every walk that
// tests segments.
__host__ __device__ __noinline__ inline int orient_exact(double p1x, double p1y, double p2x, double p2y,
                                            double cx, double cy) {
#pragma clang fp contract(off)
  const double fa[6] = {p1x, -p1y, p2x, -p2y, cx, -cy};
  const double fb[6] = {p2y, p2x, cy, cx, p1y, p1x};
  // Grow a nonoverlapping expansion h[0..12) by the twelve terms, one at a
  // time (increasing magnitude, zero components allowed). The loops have
  // constant bounds, so the expansion stays in registers.
  double h[12];
#pragma unroll
  for (int i = 0; i < 12; i++) {
    double q, err;
    two_prod(fa[i / 2], fb[i / 2], q, err);
    if (i & 1) q = err;  // term 2k: product k rounded; term 2k+1: its error
#pragma unroll
    for (int j = 0; j < i; j++) {
      double x, y;
      two_sum(q, h[j], x, y);
      h[j] = y;
      q = x;
    }
    h[i] = q;
  }
  // The most significant nonzero component carries the sign.
  int sign = 0;
#pragma unroll
  for (int i = 0; i < 12; i++)
    if (h[i] != 0.0) sign = h[i] > 0.0 ? 1 : -1;
  return sign;
}

// A segment as the tests take it: its ends s = (x1, y1, x2, y2) and the
// rounded differences dx = x2 - x1, dy = y2 - y1 (rounding keeps their signs
// exact: a difference of doubles rounds to 0 only when it is 0).
struct Segment {
  double s[4];
  double dx, dy;
};
__host__ __device__ __forceinline__ Segment make_segment(double x1, double y1, double x2,
                                                         double y2) {
  Segment g;
  g.s[0] = x1;
  g.s[1] = y1;
  g.s[2] = x2;
  g.s[3] = y2;
  g.dx = x2 - x1;
  g.dy = y2 - y1;
  return g;
}

// The corners of a box at which o(c) = orient(p1, p2, c) takes its largest
// and smallest value: o(c) = dx (cy - y1) - dy (cx - x1) grows with cy when
// dx > 0 and falls with cx when dy > 0 (the choice is exact, as the signs of
// dx and dy are). A closed segment and a closed box whose bounding boxes
// overlap meet iff o(max corner) >= 0 and o(min corner) <= 0.
struct SegmentCorners {
  double max_x, max_y, min_x, min_y;
};
__host__ __device__ __forceinline__ SegmentCorners segment_corners(const Segment& g, double bx0,
                                                                   double by0, double bx1,
                                                                   double by1) {
  SegmentCorners c;
  c.max_y = g.dx >= 0.0 ? by1 : by0;
  c.min_y = g.dx >= 0.0 ? by0 : by1;
  c.max_x = g.dy >= 0.0 ? bx0 : bx1;
  c.min_x = g.dy >= 0.0 ? bx1 : bx0;
  return c;
}

// o(c) in double with its rounding-error bound: 1 or -1 when the sign is
// certain, 0 when it is not.
__host__ __device__ __forceinline__ int orient_filtered(const Segment& g, double cx, double cy) {
#pragma clang fp contract(off)
  const double a = g.dx * (cy - g.s[1]), b = g.dy * (cx - g.s[0]);
  const double o = a - b;
  // Two differences, a product and the final difference round once each per
  // term: |error| <= (4 u + O(u^2)) (|a| + |b|), u = 2^-53; the bound keeps a
  // margin over that, and an absolute floor covers products in the
  // subnormal range.
  const double bound = 1.0e-15 * (fabs(a) + fabs(b)) + 1.0e-300;
  if (o > bound) return 1;
  if (o < -bound) return -1;
  return 0;
}

// Bounding boxes of the segment and the box (x0, y0, x1, y1) overlap
// (closed; false for a NaN box).
__host__ __device__ __forceinline__ bool segment_bbox_overlap(const Segment& g, double bx0,
                                                              double by0, double bx1, double by1) {
  const double sx0 = fmin(g.s[0], g.s[2]), sx1 = fmax(g.s[0], g.s[2]);
  const double sy0 = fmin(g.s[1], g.s[3]), sy1 = fmax(g.s[1], g.s[3]);
  return sx0 <= bx1 && sx1 >= bx0 && sy0 <= by1 && sy1 >= by0;
}

// Exact, for a box whose bounding-box overlap with the segment is already
// known: the box's corners are not all strictly on one side of the line.
__host__ __device__ inline bool segment_line_crosses(const Segment& g, double bx0, double by0,
                                                     double bx1, double by1) {
  const SegmentCorners c = segment_corners(g, bx0, by0, bx1, by1);
  int hi = orient_filtered(g, c.max_x, c.max_y);
  if (hi == 0) hi = orient_exact(g.s[0], g.s[1], g.s[2], g.s[3], c.max_x, c.max_y);
  if (hi < 0) return false;
  int lo = orient_filtered(g, c.min_x, c.min_y);
  if (lo == 0) lo = orient_exact(g.s[0], g.s[1], g.s[2], g.s[3], c.min_x, c.min_y);
  return lo <= 0;
}

// Necessary condition for the line through the segment to meet the box or
// anything inside it.
__host__ __device__ __forceinline__ bool segment_line_may_cross(const Segment& g, double bx0,
                                                                double by0, double bx1,
                                                                double by1) {
  const SegmentCorners c = segment_corners(g, bx0, by0, bx1, by1);
  return orient_filtered(g, c.max_x, c.max_y) >= 0 && orient_filtered(g, c.min_x, c.min_y) <= 0;
}

// Exact: the closed segment meets the closed box.
__host__ __device__ inline bool segment_meets_box(const Segment& g, double bx0, double by0,
                                                  double bx1, double by1) {
  return segment_bbox_overlap(g, bx0, by0, bx1, by1) &&
         segment_line_crosses(g, bx0, by0, bx1, by1);
}

// ---- float segments against float boxes (a float index) ----------------
// The same two corners evaluated in float first: o = dx (cy - y1) - dy (cx -
// x1) with each operation rounded once, |error| <= (4 u + O(u^2)) (|a| + |b|)
// for u = 2^-24; the bound keeps a margin over that and an absolute floor
// covers underflow. Only a corner within the bound goes on to the double and
// exact evaluation.
struct SegmentF {
  float s[4];
  float dx, dy;
};
__host__ __device__ __forceinline__ SegmentF make_segment_f(float x1, float y1, float x2,
                                                            float y2) {
  SegmentF g;
  g.s[0] = x1;
  g.s[1] = y1;
  g.s[2] = x2;
  g.s[3] = y2;
  g.dx = x2 - x1;
  g.dy = y2 - y1;
  return g;
}
__host__ __device__ __forceinline__ int orient_filtered_f(const SegmentF& g, float cx, float cy) {
#pragma clang fp contract(off)
  const float a = g.dx * (cy - g.s[1]), b = g.dy * (cx - g.s[0]);
  const float o = a - b;
  const float bound = 1.0e-6f * (fabsf(a) + fabsf(b)) + 1.0e-37f;
  if (o > bound) return 1;
  if (o < -bound) return -1;
  return 0;
}
// The double and exact evaluation of a float segment, out of line.
__host__ __device__ __noinline__ inline bool segment_line_crosses_slow(float x1, float y1,
                                                                       float x2, float y2,
                                                                       float bx0, float by0,
                                                                       float bx1, float by1) {
  return segment_line_crosses(make_segment(x1, y1, x2, y2), bx0, by0, bx1, by1);
}
// segment_line_crosses for a float segment and a float box.
__host__ __device__ __forceinline__ bool segment_line_crosses_f(const SegmentF& g, float bx0,
                                                                float by0, float bx1, float by1) {
  const float max_y = g.dx >= 0.0f ? by1 : by0, min_y = g.dx >= 0.0f ? by0 : by1;
  const float max_x = g.dy >= 0.0f ? bx0 : bx1, min_x = g.dy >= 0.0f ? bx1 : bx0;
  const int hi = orient_filtered_f(g, max_x, max_y);
  const int lo = orient_filtered_f(g, min_x, min_y);
  if (hi < 0 || lo > 0) return false;
  if (hi > 0 && lo < 0) return true;
  return segment_line_crosses_slow(g.s[0], g.s[1], g.s[2], g.s[3], bx0, by0, bx1, by1);
}
// segment_line_may_cross for a float segment and a float box.
__host__ __device__ __forceinline__ bool segment_line_may_cross_f(const SegmentF& g, float bx0,
                                                                  float by0, float bx1,
                                                                  float by1) {
  const float max_y = g.dx >= 0.0f ? by1 : by0, min_y = g.dx >= 0.0f ? by0 : by1;
  const float max_x = g.dy >= 0.0f ? bx0 : bx1, min_x = g.dy >= 0.0f ? bx1 : bx0;
  return orient_filtered_f(g, max_x, max_y) >= 0 && orient_filtered_f(g, min_x, min_y) <= 0;
}

}  // namespace rts
