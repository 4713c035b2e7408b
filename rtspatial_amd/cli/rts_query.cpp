// Command-line front end (reference examples/spatial_index.cu:9-109 with the
// flags of examples/flags.cpp:3-11): loads a WKT box dataset, builds the
// index, runs one query batch from a WKT box or point file and prints the
// load and query times and the result count in the reference's format.
//
//   rts_query -box=boxes.wkt -box_query=queries.wkt -predicate=intersects
//   rts_query -box=boxes.wkt -point_query=points.wkt
//
// Flags take -name=value, --name=value or -name value. -parallelism is
// accepted and ignored (it sets the reference's number of OptiX BVHs);
// -predicate also accepts "within" (every data box inside the query).
#include <hip/hip_runtime_api.h>

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <limits>
#include <map>
#include <string>
#include <vector>

#include "rtspatial/io/wkt_loader.h"
#include "rtspatial/rtspatial.h"

namespace {

struct Flags {
  std::map<std::string, std::string> v = {
      {"box", ""},          {"point_query", ""}, {"box_query", ""},   {"load_factor", "0.1"},
      {"predicate", ""},    {"limit_box", "-1"}, {"limit_query", "-1"}, {"parallelism", "1"},
      {"serialize", ""}};
  const char* help[9][2] = {{"box", "path of the box dataset (wkt)"},
                            {"point_query", "path of a point query dataset (wkt)"},
                            {"box_query", "path of a box query dataset (wkt)"},
                            {"load_factor", "result queue capacity = boxes x queries x load_factor"},
                            {"predicate", "contains, intersects or within (box queries)"},
                            {"limit_box", "read at most this many boxes (-1: all)"},
                            {"limit_query", "read at most this many queries (-1: all)"},
                            {"parallelism", "accepted for compatibility; ignored"},
                            {"serialize", "a directory for a binary copy of the parsed boxes"}};

  bool parse(int argc, char** argv) {
    for (int i = 1; i < argc; i++) {
      std::string a = argv[i];
      size_t k = a.find_first_not_of('-');
      if (k == 0 || k == std::string::npos) return bad(a);
      a = a.substr(k);
      std::string name = a, value;
      const size_t eq = a.find('=');
      if (eq != std::string::npos) {
        name = a.substr(0, eq);
        value = a.substr(eq + 1);
      } else if (i + 1 < argc) {
        value = argv[++i];
      } else {
        return bad(a);
      }
      if (!v.count(name)) return bad(name);
      v[name] = value;
    }
    return true;
  }
  bool bad(const std::string& what) {
    std::cerr << "unknown or incomplete flag: " << what << "\n";
    return false;
  }
  void usage(const char* prog) const {
    std::cerr << "Usage: " << prog << " -box=<wkt> (-box_query=<wkt> -predicate=<p> | "
                                      "-point_query=<wkt>) [flags]\n";
    for (const auto& h : help) std::cerr << "  -" << h[0] << ": " << h[1] << "\n";
  }
  int i(const std::string& k) const { return std::atoi(v.at(k).c_str()); }
  double d(const std::string& k) const { return std::atof(v.at(k).c_str()); }
  const std::string& s(const std::string& k) const { return v.at(k); }
};

double ms_since(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace

int main(int argc, char** argv) {
  Flags f;
  if (argc == 1) {
    f.usage(argv[0]);
    return 1;
  }
  if (!f.parse(argc, argv)) {
    f.usage(argv[0]);
    return 1;
  }
  using namespace rtspatial;
  using box_t = Envelope<Point<float, 2>>;
  using point_t = Point<float, 2>;
  const int limit_box = f.i("limit_box") > 0 ? f.i("limit_box") : std::numeric_limits<int>::max();
  const int limit_query =
      f.i("limit_query") > 0 ? f.i("limit_query") : std::numeric_limits<int>::max();
  try {
    if (f.s("box").empty()) throw Exception("-box is required");
    std::vector<box_t> boxes = LoadBoxes(f.s("box"), f.s("serialize"), limit_box);
    std::cout << "Loaded boxes " << boxes.size() << std::endl;
    device_vector<box_t> d_boxes(boxes);

    SpatialIndex<float, 2> index;
    Config config;
    config.max_geometries = boxes.size();
    Stream stream;
    index.Init(config);
    auto t0 = std::chrono::steady_clock::now();
    index.Insert(ArrayView<box_t>(d_boxes.data(), d_boxes.size()), stream.hip_stream());
    index.Build(stream.hip_stream());  // the reference builds inside Insert
    stream.Sync();
    const double t_load = ms_since(t0);

    double t_query = 0;
    size_t n_results = 0;
    Queue results;
    if (!f.s("box_query").empty()) {
      std::vector<box_t> queries = LoadBoxes(f.s("box_query"), limit_query);
      results.Init(std::max<size_t>(1, (size_t)(boxes.size() * queries.size() * f.d("load_factor"))));
      device_vector<box_t> d_queries(queries);
      std::cout << "Loaded box queries " << queries.size() << std::endl;
      Predicate p;
      const std::string& pred = f.s("predicate");
      if (pred == "contains") p = Predicate::kContains;
      else if (pred == "intersects") p = Predicate::kIntersects;
      else if (pred == "within") p = Predicate::kWithin;
      else throw Exception("Unsupported predicate: " + pred);
      t0 = std::chrono::steady_clock::now();
      index.Query(p, ArrayView<box_t>(d_queries.data(), d_queries.size()), RTS_HANDLER_COLLECT,
                  results.DeviceObject(), stream.hip_stream());
      n_results = results.size(stream.hip_stream());
      t_query = ms_since(t0);
    } else if (!f.s("point_query").empty()) {
      std::vector<point_t> queries = LoadPoints(f.s("point_query"), limit_query);
      results.Init(std::max<size_t>(1, (size_t)(boxes.size() * queries.size() * f.d("load_factor"))));
      device_vector<point_t> d_queries(queries);
      std::cout << "Loaded point queries " << queries.size() << std::endl;
      t0 = std::chrono::steady_clock::now();
      index.Query(Predicate::kContains, ArrayView<point_t>(d_queries.data(), d_queries.size()),
                  RTS_HANDLER_COLLECT, results.DeviceObject(), stream.hip_stream());
      n_results = results.size(stream.hip_stream());
      t_query = ms_since(t0);
    }
    index.Sync(stream.hip_stream());
    if (n_results > results.capacity())
      std::cerr << "note: " << n_results << " results, queue holds " << results.capacity()
                << " (raise -load_factor to keep them all)\n";
    // The reference's output line (spatial_index.cu:105-106).
    std::cout << "RT, load " << t_load << " ms, query " << t_query << " ms, results: " << n_results
              << std::endl;
  } catch (const std::exception& e) {
    std::cerr << "error: " << e.what() << std::endl;
    return 2;
  }
  return 0;
}
