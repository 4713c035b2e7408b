"""Runs the C++ facade test binary (tests/cpp/test_spatial_index.cpp, the
reference's gtest suite restated against include/rtspatial/) on the GPU."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.parametrize("name", ["test_spatial_index", "test_user_handlers"])
def test_cpp_facade_suite(name):
    # test_user_handlers is compiled as HIP: the reference's handler functions
    # and handler functors of its own, instantiated into the traversal kernels.
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    exe = os.path.join(ROOT, "tests", "cpp", "bin", name)
    assert os.path.exists(exe), "build first: python -c 'import __graft_entry__ as g; g.build()'"
    print(f"[gpu-test] C++ suite {name}", flush=True)
    r = subprocess.run([exe], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout
    assert "0 failed" in r.stdout


CLI = os.path.join(ROOT, "rtspatial_amd", "bin", "rts_query")


@pytest.mark.gpu
@pytest.mark.parametrize("query", ["intersects", "contains", "within", "points"])
def test_cli_counts_match_oracle(oracle, tmp_path, query):
    # The command-line front end over generated WKT files: its result count
    # equals the oracle's for the same float32 inputs.
    from rtspatial_amd import datagen, wkt
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    assert os.path.exists(CLI), "build first: python -c 'import __graft_entry__ as g; g.build()'"
    box_path = tmp_path / "boxes.wkt"
    with open(box_path, "w") as f:
        datagen.generate(f, "gaussian", 20000, geometry="box", maxsize=[0.01, 0.01], seed=11)
    q_path = tmp_path / "queries.wkt"
    with open(q_path, "w") as f:
        if query == "points":
            datagen.generate(f, "uniform", 5000, geometry="point", seed=12)
        else:
            datagen.generate(f, "uniform", 2000, geometry="box", maxsize=[0.03, 0.03], seed=12)
    boxes = wkt.load_boxes(str(box_path))
    if query == "points":
        args = [f"-point_query={q_path}"]
        want = oracle.bf_count(oracle.CONTAINS, boxes, wkt.load_points(str(q_path)))
    else:
        args = [f"-box_query={q_path}", f"-predicate={query}"]
        pred = {"intersects": oracle.INTERSECTS, "contains": oracle.CONTAINS,
                "within": oracle.WITHIN}[query]
        want = oracle.bf_count(pred, boxes, wkt.load_boxes(str(q_path)))
    print(f"[gpu-test] cli {query}", flush=True)
    r = subprocess.run([CLI, f"-box={box_path}", *args, "-load_factor=0.01"],
                       stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert f"Loaded boxes {boxes.shape[0]}" in r.stdout
    last = r.stdout.strip().splitlines()[-1]
    assert last.startswith("RT, load ") and last.endswith(f"results: {want}"), last
