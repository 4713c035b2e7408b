"""Runs the C++ facade test binary (tests/cpp/test_spatial_index.cpp, the
reference's gtest suite restated against include/rtspatial/) on the GPU."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "bin", "test_spatial_index")


def test_cpp_facade_suite():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    assert os.path.exists(BIN), "build first: python -c 'import __graft_entry__ as g; g.build()'"
    r = subprocess.run([BIN], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout
    assert "0 failed" in r.stdout
