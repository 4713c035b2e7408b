// Host build of the segment-box predicates the traversal kernels use
// (include/rtspatial/device/segment.h), for checks against the oracle on the
// CPU. Reads records from stdin and writes one line of flags per record.
//   f32 record: 8 floats  (x1, y1, x2, y2, bx0, by0, bx1, by1)
//   f64 record: 8 doubles
// Flags: exact meets, node filter (may meet), and for f32 the float-filter
// path (bounding-box test + segment_line_crosses_f / segment_line_may_cross_f).
#include <cstdio>
#include <cstring>

#include "rtspatial/device/segment.h"

int main(int argc, char** argv) {
  const bool f64 = argc > 1 && std::strcmp(argv[1], "f64") == 0;
  for (;;) {
    double v[8];
    if (f64) {
      if (std::fread(v, sizeof(double), 8, stdin) != 8) break;
    } else {
      float f[8];
      if (std::fread(f, sizeof(float), 8, stdin) != 8) break;
      for (int k = 0; k < 8; k++) v[k] = f[k];
    }
    const rts::Segment g = rts::make_segment(v[0], v[1], v[2], v[3]);
    const bool bb = rts::segment_bbox_overlap(g, v[4], v[5], v[6], v[7]);
    const int exact = rts::segment_meets_box(g, v[4], v[5], v[6], v[7]);
    const int may = bb && rts::segment_line_may_cross(g, v[4], v[5], v[6], v[7]);
    int fexact = exact, fmay = may;
    if (!f64) {
      const rts::SegmentF h = rts::make_segment_f((float)v[0], (float)v[1], (float)v[2],
                                                  (float)v[3]);
      fexact = bb && rts::segment_line_crosses_f(h, (float)v[4], (float)v[5], (float)v[6],
                                                 (float)v[7]);
      fmay = bb && rts::segment_line_may_cross_f(h, (float)v[4], (float)v[5], (float)v[6],
                                                 (float)v[7]);
    }
    std::printf("%d %d %d %d\n", exact, may, fexact, fmay);
  }
  return 0;
}
