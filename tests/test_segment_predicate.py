"""The traversal's segment-box predicates (include/rtspatial/device/segment.h),
compiled for the host, against the oracle's exact segment test on the CPU:
the exact predicate (double filter + expansion fallback), the float-filter
path of float indexes, and the node filters, which must never reject a box
the segment meets."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from test_oracle import _adversarial_segments

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def seg_host(tmp_path_factory):
    gxx = shutil.which("g++")
    if not gxx or not os.path.isdir("/opt/rocm/include"):
        pytest.skip("no g++ or ROCm headers")
    out = str(tmp_path_factory.mktemp("seg") / "segment_host")
    subprocess.run([gxx, "-O2", "-std=c++17", "-ffp-contract=off", "-D__HIP_PLATFORM_AMD__",
                    f"-I{ROOT}/include", "-I/opt/rocm/include",
                    os.path.join(ROOT, "tests", "cpp", "segment_host.cpp"), "-o", out],
                   check=True)
    return out


def run(exe, segs, boxes, dtype):
    rec = np.ascontiguousarray(np.concatenate([segs, boxes], axis=1).astype(dtype))
    res = subprocess.run([exe, "f64" if dtype == np.float64 else "f32"], input=rec.tobytes(),
                         stdout=subprocess.PIPE, check=True)
    return np.array([list(map(int, line.split())) for line in res.stdout.decode().splitlines()])


def cases(dtype):
    rng = np.random.default_rng(31 if dtype == np.float32 else 32)
    pairs = _adversarial_segments(rng, 4000, dtype)
    grid_s = (rng.integers(0, 9, (3000, 4)) / 8).astype(dtype)
    lo = rng.integers(0, 9, (3000, 2))
    hi = lo + rng.integers(0, 3, (3000, 2))
    grid_b = (np.concatenate([lo, hi], axis=1) / 8).astype(dtype)
    segs = np.concatenate([np.array([p[0] for p in pairs]), grid_s])
    boxes = np.concatenate([np.array([p[1] for p in pairs]), grid_b])
    return segs, boxes


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_device_segment_predicates_match_oracle(oracle, seg_host, dtype):
    segs, boxes = cases(dtype)
    got = run(seg_host, segs, boxes, dtype)
    ev = oracle.evaluate if dtype == np.float32 else oracle.evaluate64
    want = np.array([ev(oracle.INTERSECTS, b, s, lines=True) for s, b in zip(segs, boxes)])
    assert 0.2 < want.mean() < 0.9
    np.testing.assert_array_equal(got[:, 0], want)       # exact predicate
    np.testing.assert_array_equal(got[:, 2], want)       # float-filter path
    assert np.all(got[want == 1, 1] == 1)                # node filters are necessary
    assert np.all(got[want == 1, 3] == 1)
    assert np.mean(got[want == 0, 1] == 0) > 0.5         # and prune most misses


@pytest.mark.parametrize("scale", [2.0 ** -400, 2.0 ** 400])
def test_device_segment_predicate_extreme_doubles(oracle, seg_host, scale):
    # The exact predicate's documented range for doubles (magnitudes 0 and
    # 2^-450 .. 2^500): adversarial cases scaled by a power of two keep their
    # answers, and the host build of the device predicate must still give
    # the exact rational one.
    segs, boxes = cases(np.float64)
    segs = segs * scale
    boxes = boxes * scale
    got = run(seg_host, segs, boxes, np.float64)
    want = np.array([oracle.segment_meets_box(s, b) for s, b in zip(segs, boxes)])
    np.testing.assert_array_equal(got[:, 0], want)
    assert np.all(got[want, 1] == 1)  # the node filter stays a necessary condition
