// WKT ingest for the spatial index (reference examples/wkt_loader.h:25-218,
// there built on Boost.Geometry; here a small parser with no dependencies).
//
// LoadBoxes: one envelope per line for POLYGON, one per member polygon for
//   MULTIPOLYGON: the min/max over the outer ring's vertices
//   (wkt_loader.h:30-43, 48-67).
// LoadPoints: POINT gives its point; POLYGON / MULTIPOLYGON give every vertex
//   of their outer rings (wkt_loader.h:76-113).
// Coordinates are parsed as float (strtof, correctly rounded), the reference's
// coord_t. Reading stops after the line on which the count reaches `limit`
// (so a MULTIPOLYGON may add past it, as in the reference). Empty lines are
// skipped; any other geometry throws rtspatial::Exception (the reference
// prints "Bad Geometry" and aborts).
//
// LoadBoxes(path, cache_dir, limit) keeps a binary copy of the parsed boxes in
// cache_dir (wkt_loader.h:150-218 does this with a Boost archive; the format
// here is this header's own: "RTSBOX01", a little-endian uint64 count, then
// count x {min x, min y, max x, max y} floats).
#ifndef RTSPATIAL_AMD_IO_WKT_LOADER_H
#define RTSPATIAL_AMD_IO_WKT_LOADER_H

#include <sys/stat.h>

#include <algorithm>
#include <cerrno>
#include <climits>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <limits>
#include <string>
#include <vector>

#include "rtspatial/geom/envelope.h"
#include "rtspatial/geom/point.h"
#include "rtspatial/utils/exception.h"

namespace rtspatial {
namespace wkt {

using point_t = Point<float, 2>;
using box_t = Envelope<point_t>;

// A cursor over one WKT line.
class Reader {
 public:
  explicit Reader(const std::string& s) : p_(s.c_str()), end_(s.c_str() + s.size()), line_(s) {}

  void skip_ws() {
    while (p_ < end_ && (*p_ == ' ' || *p_ == '\t' || *p_ == '\r' || *p_ == '\n')) ++p_;
  }
  bool peek(char c) {
    skip_ws();
    return p_ < end_ && *p_ == c;
  }
  void expect(char c) {
    if (!peek(c)) fail(std::string("expected '") + c + "'");
    ++p_;
  }
  bool keyword(const char* kw) {
    skip_ws();
    const size_t n = std::strlen(kw);
    if ((size_t)(end_ - p_) >= n && std::strncmp(p_, kw, n) == 0) {
      p_ += n;
      return true;
    }
    return false;
  }
  float number() {
    skip_ws();
    char* e = nullptr;
    errno = 0;
    const float v = std::strtof(p_, &e);
    if (e == p_ || e > end_) fail("expected a number");
    p_ = e;
    return v;
  }
  point_t point() {
    const float x = number();
    const float y = number();
    while (p_ < end_ && *p_ != ',' && *p_ != ')') {  // extra coordinates (Z, M) are ignored
      skip_ws();
      if (p_ < end_ && *p_ != ',' && *p_ != ')') number();
    }
    return point_t(x, y);
  }
  // "(x y, x y, ...)"
  std::vector<point_t> ring() {
    std::vector<point_t> pts;
    expect('(');
    pts.push_back(point());
    while (peek(',')) {
      ++p_;
      pts.push_back(point());
    }
    expect(')');
    return pts;
  }
  // "((outer), (hole), ...)": the outer ring; holes are read and dropped.
  std::vector<point_t> polygon_outer() {
    expect('(');
    std::vector<point_t> outer = ring();
    while (peek(',')) {
      ++p_;
      (void)ring();
    }
    expect(')');
    return outer;
  }
  [[noreturn]] void fail(const std::string& why) const {
    throw Exception("Bad Geometry (" + why + "): " + line_.substr(0, 120));
  }

 private:
  const char* p_;
  const char* end_;
  const std::string& line_;
};

inline box_t ring_box(const std::vector<point_t>& pts) {
  float lo[2] = {std::numeric_limits<float>::max(), std::numeric_limits<float>::max()};
  float hi[2] = {std::numeric_limits<float>::lowest(), std::numeric_limits<float>::lowest()};
  for (const point_t& p : pts) {
    lo[0] = std::min(lo[0], p.get_x());
    hi[0] = std::max(hi[0], p.get_x());
    lo[1] = std::min(lo[1], p.get_y());
    hi[1] = std::max(hi[1], p.get_y());
  }
  return box_t(point_t(lo[0], lo[1]), point_t(hi[0], hi[1]));
}

// Calls on_ring(outer ring) for each polygon of a POLYGON / MULTIPOLYGON
// line, or on_point(p) for a POINT line (if on_point is allowed).
template <typename OnRing, typename OnPoint>
void parse_line(const std::string& line, bool points_ok, OnRing on_ring, OnPoint on_point) {
  Reader r(line);
  if (r.keyword("MULTIPOLYGON")) {
    if (r.keyword("EMPTY")) return;
    r.expect('(');
    on_ring(r.polygon_outer());
    while (r.peek(',')) {
      r.expect(',');
      on_ring(r.polygon_outer());
    }
    r.expect(')');
  } else if (r.keyword("POLYGON")) {
    if (r.keyword("EMPTY")) return;
    on_ring(r.polygon_outer());
  } else if (points_ok && r.keyword("POINT")) {
    r.expect('(');
    on_point(r.point());
    r.expect(')');
  } else {
    r.fail("unsupported geometry");
  }
}

inline bool blank(const std::string& s) {
  return s.find_first_not_of(" \t\r\n") == std::string::npos;
}

inline std::ifstream open_or_throw(const std::string& path) {
  std::ifstream ifs(path);
  if (!ifs) throw Exception("cannot open " + path);
  return ifs;
}

}  // namespace wkt

inline std::vector<wkt::box_t> LoadBoxes(const std::string& path,
                                         int limit = std::numeric_limits<int>::max()) {
  std::ifstream ifs = wkt::open_or_throw(path);
  std::vector<wkt::box_t> boxes;
  std::string line;
  while (std::getline(ifs, line)) {
    if (wkt::blank(line)) continue;
    wkt::parse_line(
        line, false, [&](const std::vector<wkt::point_t>& ring) { boxes.push_back(wkt::ring_box(ring)); },
        [](const wkt::point_t&) {});
    if (boxes.size() >= (size_t)limit) break;
  }
  return boxes;
}

inline std::vector<wkt::point_t> LoadPoints(const std::string& path,
                                            int limit = std::numeric_limits<int>::max()) {
  std::ifstream ifs = wkt::open_or_throw(path);
  std::vector<wkt::point_t> points;
  std::string line;
  while (std::getline(ifs, line)) {
    if (wkt::blank(line)) continue;
    wkt::parse_line(
        line, true,
        [&](const std::vector<wkt::point_t>& ring) {
          points.insert(points.end(), ring.begin(), ring.end());
        },
        [&](const wkt::point_t& p) { points.push_back(p); });
    if (points.size() >= (size_t)limit) break;
  }
  return points;
}

inline void SerializeBoxes(const std::string& file, const std::vector<wkt::box_t>& boxes) {
  std::ofstream ofs(file, std::ios::binary);
  if (!ofs) throw Exception("cannot write " + file);
  const uint64_t n = boxes.size();
  ofs.write("RTSBOX01", 8);
  ofs.write(reinterpret_cast<const char*>(&n), sizeof(n));
  ofs.write(reinterpret_cast<const char*>(boxes.data()), (std::streamsize)(n * sizeof(wkt::box_t)));
}

inline std::vector<wkt::box_t> DeserializeBoxes(const std::string& file) {
  std::ifstream ifs(file, std::ios::binary);
  char magic[8] = {};
  uint64_t n = 0;
  if (!ifs.read(magic, 8) || std::memcmp(magic, "RTSBOX01", 8) != 0 ||
      !ifs.read(reinterpret_cast<char*>(&n), sizeof(n)))
    throw Exception("not a box cache: " + file);
  std::vector<wkt::box_t> boxes(n);
  if (!ifs.read(reinterpret_cast<char*>(boxes.data()), (std::streamsize)(n * sizeof(wkt::box_t))))
    throw Exception("truncated box cache: " + file);
  return boxes;
}

// wkt_loader.h:150-218: with a cache directory, the boxes of (path, limit)
// are read from the cache if present, else parsed and written there.
inline std::vector<wkt::box_t> LoadBoxes(const std::string& path, const std::string& cache_dir,
                                         int limit = std::numeric_limits<int>::max()) {
  if (cache_dir.empty()) return LoadBoxes(path, limit);
  struct stat st;
  if (stat(cache_dir.c_str(), &st) != 0) {
    if (mkdir(cache_dir.c_str(), 0755) != 0 && errno != EEXIST)
      throw Exception("cannot create " + cache_dir);
  } else if (!S_ISDIR(st.st_mode)) {
    throw Exception(cache_dir + " is not a directory");
  }
  std::string escaped = path;
  std::replace(escaped.begin(), escaped.end(), '/', '-');
  const std::string cache = cache_dir + "/" + escaped + "_limit_" + std::to_string(limit) + ".bin";
  if (stat(cache.c_str(), &st) == 0) return DeserializeBoxes(cache);
  std::vector<wkt::box_t> boxes = LoadBoxes(path, limit);
  SerializeBoxes(cache, boxes);
  return boxes;
}

}  // namespace rtspatial
#endif
