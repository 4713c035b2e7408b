"""WKT ingest (include/rtspatial/io/wkt_loader.h and rtspatial_amd/wkt.py,
mirroring the reference's examples/wkt_loader.h), the dataset generator
(rtspatial_amd/datagen.py, mirroring dataset/generator.py) and the
command-line front end (rtspatial_amd/bin/rts_query, mirroring
examples/spatial_index.cu)."""
import io
import math
import os
import shutil
import subprocess

import numpy as np
import pytest

from rtspatial_amd import datagen, wkt

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
REF_DATA = "/root/reference/dataset"

SAMPLE = """MULTIPOLYGON(((0 0, 1 0, 1 1, 0 0)), ((2 2, 3 2, 3 5, 2 2), (2.1 2.1, 2.2 2.1, 2.2 2.2, 2.1 2.1)))

POLYGON ((0.1 0.2, 0.3 0.4, 0.05 0.9, 0.1 0.2), (0.11 0.21, 0.12 0.21, 0.12 0.22, 0.11 0.21))
POLYGON((-1.5e-1 2.5E+0,3 4,5 -6,-1.5e-1 2.5E+0))
MULTIPOLYGON EMPTY
"""


def write(tmp_path, name, text):
    p = tmp_path / name
    p.write_text(text)
    return str(p)


def test_python_loader_boxes_and_points(tmp_path):
    path = write(tmp_path, "s.wkt", SAMPLE)
    b = wkt.load_boxes(path)
    want = np.float32([[0, 0, 1, 1], [2, 2, 3, 5], [0.05, 0.2, 0.3, 0.9], [-0.15, -6, 5, 4]])
    assert b.dtype == np.float32 and np.array_equal(b, want)
    p = wkt.load_points(write(tmp_path, "p.wkt", SAMPLE + "POINT (7 8)\nPOINT(1e-3 -2)\n"))
    assert p.shape == (4 + 4 + 4 + 4 + 2, 2)  # outer rings only, holes dropped
    assert np.array_equal(p[-2:], np.float32([[7, 8], [1e-3, -2]]))
    # limit: reading stops after the line that reaches it (a multipolygon may overshoot)
    assert wkt.load_boxes(path, limit=1).shape[0] == 2
    assert wkt.load_boxes(path, limit=3).shape[0] == 3


def test_python_loader_rejects_other_geometries(tmp_path):
    with pytest.raises(ValueError, match="Bad Geometry"):
        wkt.load_boxes(write(tmp_path, "bad.wkt", "LINESTRING(0 0, 1 1)\n"))
    with pytest.raises(ValueError, match="Bad Geometry"):
        wkt.load_boxes(write(tmp_path, "pt.wkt", "POINT(1 2)\n"))  # points are not boxes


def test_float32_rounding_is_from_decimal_text(tmp_path):
    # 0.1000000014901161193847656249 rounds down to float32 0.1 but rounds up if
    # it is first rounded to double (double rounding): the loaders parse
    # straight to float32, as the reference's coord_t = float does.
    tok = "0.10000000149011611938476562"
    b = wkt.load_boxes(write(tmp_path, "r.wkt", f"POLYGON(({tok} 0, 1 0, 1 1, {tok} 0))\n"))
    assert b[0, 0] == np.float32(0.1)


@pytest.mark.skipif(not os.path.isdir(REF_DATA), reason="the reference's dataset files are not here")
def test_reference_wkt_files_match_golden_fixture():
    g = np.load(os.path.join(GOLDEN, "wkt_1000.npz"))
    for kind in ("uniform", "gaussian"):
        assert np.array_equal(wkt.load_boxes(f"{REF_DATA}/box/{kind}_1000.wkt"), g[f"{kind}_boxes"])
        assert np.array_equal(wkt.load_points(f"{REF_DATA}/point/{kind}_1000.wkt"),
                              g[f"{kind}_points"])


CPP_DUMP = r"""
#include <cstdio>
#include <string>
#include "rtspatial/io/wkt_loader.h"
int main(int argc, char** argv) {
  std::string mode = argv[1], path = argv[2];
  int limit = argc > 3 ? std::atoi(argv[3]) : 2147483647;
  try {
    if (mode == "boxes" || mode == "cached") {
      auto b = mode == "boxes" ? rtspatial::LoadBoxes(path, limit)
                               : rtspatial::LoadBoxes(path, std::string(argv[4]), limit);
      std::fwrite(b.data(), sizeof(b[0]), b.size(), stdout);
    } else {
      auto p = rtspatial::LoadPoints(path, limit);
      std::fwrite(p.data(), sizeof(p[0]), p.size(), stdout);
    }
  } catch (const std::exception& e) {
    std::fprintf(stderr, "%s\n", e.what());
    return 3;
  }
  return 0;
}
"""


@pytest.fixture(scope="module")
def cpp_dump(tmp_path_factory):
    gxx = shutil.which("g++")
    if not gxx or not os.path.exists("/opt/rocm/include/hip/hip_runtime_api.h"):
        pytest.skip("no g++ or HIP headers")
    d = tmp_path_factory.mktemp("wktcpp")
    src = d / "dump.cpp"
    src.write_text(CPP_DUMP)
    exe = d / "dump"
    subprocess.run([gxx, "-O1", "-std=c++17", f"-I{ROOT}/include", "-I/opt/rocm/include",
                    "-D__HIP_PLATFORM_AMD__", str(src), "-o", str(exe)], check=True)
    return str(exe)


def run_dump(exe, *args):
    r = subprocess.run([exe, *args], stdout=subprocess.PIPE, stderr=subprocess.PIPE)
    return r.returncode, r.stdout, r.stderr.decode()


def test_cpp_loader_matches_python_loader(cpp_dump, tmp_path):
    path = write(tmp_path, "s.wkt", SAMPLE + "POINT (7 8)\n")
    boxes_path = write(tmp_path, "b.wkt", SAMPLE)
    rc, out, _ = run_dump(cpp_dump, "boxes", boxes_path)
    assert rc == 0 and np.array_equal(np.frombuffer(out, np.float32).reshape(-1, 4),
                                      wkt.load_boxes(boxes_path))
    rc, out, _ = run_dump(cpp_dump, "points", path)
    assert rc == 0 and np.array_equal(np.frombuffer(out, np.float32).reshape(-1, 2),
                                      wkt.load_points(path))
    rc, out, _ = run_dump(cpp_dump, "boxes", boxes_path, "3")
    assert rc == 0 and np.frombuffer(out, np.float32).reshape(-1, 4).shape[0] == 3
    rc, _, err = run_dump(cpp_dump, "boxes", write(tmp_path, "bad.wkt", "LINESTRING(0 0,1 1)\n"))
    assert rc == 3 and "Bad Geometry" in err


def test_cpp_loader_box_cache_round_trip(cpp_dump, tmp_path):
    src = write(tmp_path, "g.wkt", "".join(
        datagen.wkt_box(lo, hi) + "\n"
        for lo, hi in datagen.boxes("gaussian", 500, 2, datagen.Draw(5), maxsize=[0.01, 0.01])))
    cache = str(tmp_path / "cache")
    rc, a, _ = run_dump(cpp_dump, "cached", src, "400", cache)
    assert rc == 0 and len(os.listdir(cache)) == 1
    os.remove(src)  # the second load must come from the cache
    rc, b, _ = run_dump(cpp_dump, "cached", src, "400", cache)
    assert rc == 0 and a == b and len(a) == 400 * 16


@pytest.mark.skipif(not os.path.isdir(REF_DATA), reason="the reference's dataset files are not here")
def test_cpp_loader_reads_reference_files_like_golden_fixture(cpp_dump):
    g = np.load(os.path.join(GOLDEN, "wkt_1000.npz"))
    for kind in ("uniform", "gaussian"):
        rc, out, _ = run_dump(cpp_dump, "boxes", f"{REF_DATA}/box/{kind}_1000.wkt")
        assert rc == 0 and np.array_equal(np.frombuffer(out, np.float32).reshape(-1, 4),
                                          g[f"{kind}_boxes"])
        rc, out, _ = run_dump(cpp_dump, "points", f"{REF_DATA}/point/{kind}_1000.wkt")
        assert rc == 0 and np.array_equal(np.frombuffer(out, np.float32).reshape(-1, 2),
                                          g[f"{kind}_points"])


# ---------------------------------------------------------------- datagen --
def gen(**kw):
    buf = io.StringIO()
    datagen.generate(buf, **kw)
    return buf.getvalue().splitlines()


def test_generator_is_seeded_and_writes_wkt_the_loader_reads(tmp_path):
    a = gen(distribution="uniform", cardinality=300, geometry="box", maxsize=[0.005, 0.005], seed=7)
    b = gen(distribution="uniform", cardinality=300, geometry="box", maxsize=[0.005, 0.005], seed=7)
    c = gen(distribution="uniform", cardinality=300, geometry="box", maxsize=[0.005, 0.005], seed=8)
    assert a == b and a != c and len(a) == 300
    assert all(l.startswith("POLYGON((") for l in a)
    boxes = wkt.load_boxes(write(tmp_path, "u.wkt", "\n".join(a) + "\n"))
    w, h = boxes[:, 2] - boxes[:, 0], boxes[:, 3] - boxes[:, 1]
    assert boxes.shape == (300, 4) and (w >= 0).all() and (w <= 0.0100001).all() and (h <= 0.0100001).all()
    pts = gen(distribution="gaussian", cardinality=4000, geometry="point", seed=3)
    p = wkt.load_points(write(tmp_path, "g.wkt", "\n".join(pts) + "\n"))
    assert p.shape == (4000, 2) and abs(p.mean() - 0.5) < 0.01 and abs(p.std() - 0.1) < 0.01


@pytest.mark.parametrize("dist,kw", [
    ("diagonal", {"percentage": 0.5, "buffer": 0.1}),
    ("sierpinski", {}),
    ("bit", {"probability": 0.2, "digits": 10}),
])
def test_generator_point_distributions_stay_in_unit_square(dist, kw):
    pts = [[float(v) for v in l.split(",")] for l in
           gen(distribution=dist, cardinality=2000, geometry="point", fmt="csv", seed=1, **kw)]
    a = np.asarray(pts)
    assert a.shape == (2000, 2) and (a >= 0).all() and (a <= 1).all()
    if dist == "diagonal":
        assert np.mean(a[:, 0] == a[:, 1]) > 0.4  # about half exactly on the diagonal
    if dist == "bit":
        assert np.allclose(a * 1024, np.round(a * 1024))  # 10 binary digits


def test_parcel_boxes_tile_the_unit_square():
    # Without dither the parcels partition [0,1]^2: their areas sum to one and
    # a fine grid of sample points falls in exactly one box each.
    lines = gen(distribution="parcel", cardinality=1000, fmt="csv", seed=2, srange=0.1, dither=0.0)
    b = np.asarray([[float(v) for v in l.split(",")] for l in lines])
    assert b.shape == (1000, 4)
    assert math.isclose(float(np.sum((b[:, 2] - b[:, 0]) * (b[:, 3] - b[:, 1]))), 1.0, rel_tol=1e-9)
    g = (np.arange(97) + 0.5) / 97
    xs, ys = np.meshgrid(g, g)
    inside = ((xs.ravel()[:, None] > b[:, 0]) & (xs.ravel()[:, None] < b[:, 2]) &
              (ys.ravel()[:, None] > b[:, 1]) & (ys.ravel()[:, None] < b[:, 3]))
    assert (inside.sum(axis=1) == 1).all()


def test_cli_prints_usage_without_arguments():
    exe = os.path.join(ROOT, "rtspatial_amd", "bin", "rts_query")
    if not os.path.exists(exe):
        pytest.skip("rts_query not built")
    r = subprocess.run([exe], stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)
    assert r.returncode == 1 and "-box_query" in r.stderr and "-predicate" in r.stderr
