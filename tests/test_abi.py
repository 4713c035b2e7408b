"""CPU tests of the C-ABI library: it loads, exports every entry point that
include/rts_c.h declares, and its host-only calls behave (no GPU needed)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = []
    for h in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if not h.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names += re.findall(r"^\s*(?:const\s+)?[\w\s\*]+?\b(rts_\w+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_header_declares_entry_points():
    names = declared_functions()
    for n in ("rts_index_create", "rts_index_insert", "rts_index_delete", "rts_index_update",
              "rts_index_query_points", "rts_index_query_envelopes", "rts_sort_pairs"):
        assert n in names


def test_library_exports_every_declared_symbol():
    from rtspatial_amd import _native as N
    L = N.lib()
    for n in declared_functions():
        assert hasattr(L, n), f"{n} declared in include/ but not exported"
    bound = {s[0] for s in N.SIGNATURES}
    assert set(declared_functions()) <= bound, set(declared_functions()) - bound


def test_config_defaults_match_reference():
    # spatial_index.cuh:22-37
    from rtspatial_amd import _native as N
    c = N.rts_config()
    N.lib().rts_config_default(ctypes.byref(c))
    assert c.max_geometries == 10_000_000 and c.max_queries == 1_000_000
    assert abs(c.geom_sample_rate - 0.2) < 1e-7 and abs(c.intersect_cost_weight - 0.95) < 1e-7
    assert c.max_geom_samples == 10000 and c.max_query_samples == 1000 and c.max_parallelism == 512
    assert c.leaf_size > 0


def test_struct_layouts():
    from rtspatial_amd import _native as N
    assert ctypes.sizeof(N.rts_dev_queue) == 24
    assert ctypes.sizeof(N.rts_config) % 8 == 0
    assert ctypes.sizeof(N.rts_tune_config) == 16 and ctypes.sizeof(N.rts_tune_result) == 48


LAYOUT_C = r"""
#include <stdio.h>
#include <stddef.h>
#include "rts_c.h"
#define F(T, m) printf(#T "." #m " %zu\n", offsetof(T, m));
#define S(T) printf(#T " %zu\n", sizeof(T));
int main(void) {
  S(rts_config) F(rts_config, sort_queries) F(rts_config, coord)
  S(rts_point_d) S(rts_envelope_d) S(rts_update_d) F(rts_update_d, box)
  S(rts_update) S(rts_tune_result) S(rts_serve_config) S(rts_request) S(rts_request_result)
  S(rts_index_stats) F(rts_index_stats, delta_boxes) F(rts_config, insert_delta_pct)
  S(rts_tree_view) F(rts_tree_view, n) F(rts_tree_view, coord) S(rts_query_plan)
  S(rts_memory_usage) S(rts_line) S(rts_line_d) F(rts_line, p2) F(rts_line_d, p2)
  printf("RTS_QUERY_LINES %d\n", (int)RTS_QUERY_LINES);
  return 0;
}
"""


def test_ctypes_structs_match_the_header(tmp_path):
    # The ctypes mirror of every struct has the C compiler's size and offsets.
    import shutil
    import subprocess
    from rtspatial_amd import _native as N
    gcc = shutil.which("gcc")
    if not gcc:
        pytest.skip("no gcc")
    (tmp_path / "l.c").write_text(LAYOUT_C)
    subprocess.run([gcc, "-std=c99", f"-I{ROOT}/include", str(tmp_path / "l.c"), "-o",
                    str(tmp_path / "l")], check=True)
    out = subprocess.run([str(tmp_path / "l")], check=True, stdout=subprocess.PIPE, text=True).stdout
    got = dict(line.rsplit(" ", 1) for line in out.splitlines())
    assert int(got["rts_config"]) == ctypes.sizeof(N.rts_config) == 88
    assert int(got["rts_config.coord"]) == N.rts_config.coord.offset == 76
    assert int(got["rts_point_d"]) == 16 and int(got["rts_envelope_d"]) == 32
    assert int(got["rts_update_d"]) == 40 and int(got["rts_update_d.box"]) == 8
    assert int(got["rts_update"]) == 24
    assert int(got["rts_tree_view.n"]) == N.rts_tree_view.n.offset == 56
    assert int(got["rts_index_stats"]) == ctypes.sizeof(N.rts_index_stats) == 56
    assert int(got["rts_index_stats.delta_boxes"]) == N.rts_index_stats.delta_boxes.offset == 52
    assert int(got["rts_config.insert_delta_pct"]) == N.rts_config.insert_delta_pct.offset == 80
    assert int(got["rts_tree_view.coord"]) == N.rts_tree_view.coord.offset
    assert int(got["rts_line"]) == 16 and int(got["rts_line.p2"]) == 8
    assert int(got["rts_line_d"]) == 32 and int(got["rts_line_d.p2"]) == 16
    assert int(got["RTS_QUERY_LINES"]) == N.RTS_QUERY_LINES == 2
    for name in ("rts_tune_result", "rts_serve_config", "rts_request", "rts_request_result",
                 "rts_tree_view", "rts_query_plan", "rts_memory_usage"):
        assert int(got[name]) == ctypes.sizeof(getattr(N, name)), name


def test_coord_default_is_float():
    from rtspatial_amd import _native as N
    c = N.rts_config()
    N.lib().rts_config_default(ctypes.byref(c))
    assert c.coord == N.RTS_COORD_F32 and c.insert_delta_pct == 2


def test_tune_config_defaults_and_null_args():
    from rtspatial_amd import _native as N
    c = N.rts_tune_config()
    N.lib().rts_tune_config_default(ctypes.byref(c))
    assert c.sample == 32768 and c.reps == 2 and c.apply == 1
    assert c.leaf_sizes == (1 << 2) | (1 << 3) | (1 << 4) | (1 << 5)
    r = N.rts_tune_result()
    assert N.lib().rts_index_autotune(None, 0, None, 1, 0, 1, None, ctypes.byref(r), None) == 1


def test_version_and_error_string():
    from rtspatial_amd import _native as N
    assert b"gfx950" in N.lib().rts_version()
    assert isinstance(N.lib().rts_last_error(), bytes)


def test_null_arguments_rejected_without_gpu():
    from rtspatial_amd import _native as N
    L = N.lib()
    assert L.rts_index_create(None, None) == 1
    assert L.rts_index_insert(None, None, 1, None) == 1
    assert L.rts_sort_pairs(None, None, None, None, 4, 8, None) == 1
    assert L.rts_index_query_points(None, 1, None, 1, 0, None, None) == 1


def test_create_without_gpu_reports_missing():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    from rtspatial_amd import _native as N
    h = ctypes.c_void_p()
    assert N.lib().rts_index_create(None, ctypes.byref(h)) == 4
    assert h.value is None


def test_python_api_surface():
    import rtspatial_amd as R
    for m in ("Init", "Clear", "Insert", "Delete", "Update", "Query"):
        assert hasattr(R.SpatialIndex, m)
    assert int(R.Predicate.kIntersects) == 0 and int(R.Predicate.kContains) == 1


def test_cpp_facade_headers_compile_standalone(tmp_path):
    # The C++ facade (include/rtspatial/) is host-only C++ over the C ABI: it
    # must compile with a plain host compiler and the HIP runtime headers.
    import shutil
    import subprocess
    gxx = shutil.which("g++")
    if not gxx or not os.path.exists("/opt/rocm/include/hip/hip_runtime_api.h"):
        pytest.skip("no g++ or HIP headers")
    src = tmp_path / "use_facade.cpp"
    src.write_text(
        '#include "rtspatial/default_handlers/counting_handlers.h"\n'
        '#include "rtspatial/rtspatial.h"\n'
        "int main() { rtspatial::SpatialIndex<float, 2> ix; (void)ix.size(); return 0; }\n")
    r = subprocess.run([gxx, "-std=c++17", "-fsyntax-only", "-D__HIP_PLATFORM_AMD__",
                        "-I/opt/rocm/include", f"-I{os.path.join(ROOT, 'include')}", str(src)],
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    assert r.returncode == 0, r.stdout
