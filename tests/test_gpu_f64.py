"""GPU parity of double-coordinate indexes (SpatialIndex<double, 2>) through
the C ABI (rts_index_*_d), against the oracle's exact double join.

The reference instantiates SpatialIndex<double, 2> (spatial_index.cuh:373,
428, 458 select its double modules) and tests it with
test/test_point_queries.h:92-120 (5 results). Its double Point hands float
coordinates to the predicates (geom/point.cuh:68), so its results are the
float predicates on rounded inputs; RTS_COORD_F64_AS_F32 reproduces those,
RTS_COORD_F64 (the default) is exact in double.

Inputs are mostly doubles that are not floats, and a share of the queries
puts coordinates within one float ulp of data coordinates, where the float
bounds cannot decide and the kernels read the double boxes."""
import ctypes
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():
    pytest.skip("no GPU", allow_module_level=True)

import rtspatial_amd as R  # noqa: E402
from gpu_util import collect, count, dev, log, make_index, pinned  # noqa: E402

N = R._native
P = R.Predicate
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
THREAD, WAVE = N.RTS_TRAVERSAL_THREAD, N.RTS_TRAVERSAL_WAVE


def boxes64(rng, n, side):
    lo = rng.random((n, 2))
    hi = lo + rng.random((n, 2)) * side
    return np.concatenate([lo, hi], axis=1)  # x0, y0, x1, y1


def near(rng, boxes, n):
    """Queries whose coordinates sit on, or within 1e-13 of, data box
    coordinates (far below a float ulp near 0.5, ~3e-8)."""
    src = boxes[rng.integers(0, boxes.shape[0], n)]
    eps = rng.choice([-1e-13, -1e-15, 0.0, 1e-15, 1e-13], size=(n, 4))
    q = src + eps
    grow = rng.random(n) < 0.5  # half of them also reach over to a neighbour
    q[grow, 2:] += rng.random((int(grow.sum()), 2)) * 0.01
    return q


def points_near(rng, boxes, n):
    src = boxes[rng.integers(0, boxes.shape[0], n)]
    corner = rng.integers(0, 2, (n, 2))
    p = np.where(corner == 0, src[:, :2], src[:, 2:])
    return p + rng.choice([-1e-13, 0.0, 1e-13], size=(n, 2), p=[0.3, 0.4, 0.3])


@pytest.fixture(scope="module")
def data():
    rng = np.random.default_rng(640)
    b = boxes64(rng, 60_000, 0.01)
    q = np.concatenate([boxes64(rng, 2000, 0.05), near(rng, b, 2000)])
    pts = np.concatenate([rng.random((2000, 2)), points_near(rng, b, 2000)])
    return b, q, pts


def test_f64_known_answer_pinned_host():
    # test/test_point_queries.h:92-120: pinned double inputs, count 5.
    log("f64 tiny point case")
    with open(os.path.join(GOLDEN, "known_answers.json")) as f:
        c = json.load(f)["tiny_points"]
    ix = R.SpatialIndex(torch.float64)
    ix.Init(R.Config())
    ix.Insert(pinned(np.array(c["boxes"], np.float64)))
    pts = pinned(np.array(c["points"], np.float64))
    cnt = R.SharedValue()
    cnt.set(0)
    ix.Query(P.kContains, pts, cnt)
    assert cnt.get() == 5
    pairs, n = collect(ix, P.kContains, dev(np.array(c["points"], np.float64)))
    assert n == 5 and pairs.tolist() == c["pairs_query_box"]


@pytest.mark.parametrize("mode", [THREAD, WAVE])
@pytest.mark.parametrize("pred", [P.kIntersects, P.kContains, P.kWithin])
def test_f64_envelopes_match_oracle(oracle, data, pred, mode):
    b, q, _ = data
    log(f"f64 envelopes pred={int(pred)} mode={mode}")
    want = oracle.bf_pairs64(int(pred), b, q)
    for sort in (1, 2):
        ix = make_index(b, traversal=mode, sort_queries=sort, leaf_size=8)
        assert count(ix, pred, q) == want.shape[0]
        got, n = collect(ix, pred, q)
        assert n == want.shape[0] and np.array_equal(got, want), f"sort={sort}"


@pytest.mark.parametrize("mode", [THREAD, WAVE])
def test_f64_points_match_oracle(oracle, data, mode):
    b, _, pts = data
    log(f"f64 points mode={mode}")
    want = oracle.bf_pairs64(oracle.CONTAINS, b, pts)
    ix = make_index(b, traversal=mode, leaf_size=4)
    got, n = collect(ix, P.kContains, pts)
    assert n == want.shape[0] and np.array_equal(got, want)
    assert count(ix, P.kContains, pts) == want.shape[0]


def test_f64_flips_exact_and_reference_rounding(oracle):
    # Each case is false in double and true once rounded to float: the exact
    # index reports nothing, the rounding one (the reference's behaviour)
    # reports the pair -- both equal to their oracle.
    from test_oracle import F64_FLIPS
    log("f64 rounding flips")
    for pred, box, query in F64_FLIPS:
        b = np.array([box], np.float64)
        q = np.array([query], np.float64)
        exact = make_index(b)
        assert count(exact, pred, q) == 0, (pred, box, query)
        rounded = make_index(b, coord=N.RTS_COORD_F64_AS_F32)
        assert count(rounded, pred, q) == 1, (pred, box, query)
        assert oracle.bf_count(pred, b.astype(np.float32), q.astype(np.float32)) == 1


@pytest.mark.parametrize("mode", [THREAD, WAVE])
def test_f64_as_f32_equals_float_index_on_rounded_inputs(oracle, data, mode):
    b, q, pts = data
    log(f"f64 as f32 mode={mode}")
    ix = make_index(b, coord=N.RTS_COORD_F64_AS_F32, traversal=mode)
    bf, qf, pf = b.astype(np.float32), q.astype(np.float32), pts.astype(np.float32)
    for pred in (P.kIntersects, P.kWithin):
        got, n = collect(ix, pred, q)
        assert np.array_equal(got, oracle.bf_pairs(int(pred), bf, qf))
    got, n = collect(ix, P.kContains, pts)
    assert np.array_equal(got, oracle.bf_pairs(oracle.CONTAINS, bf, pf))


def test_f64_extreme_coordinates(oracle):
    # Beyond float range (rounds to +-inf / FLT_MAX), float subnormals and
    # double subnormals, mixed with ordinary boxes.
    log("f64 extreme coordinates")
    rng = np.random.default_rng(77)
    b = boxes64(rng, 3000, 0.05)
    specials = np.array([
        [-1e300, -1e300, 1e300, 1e300],
        [1e-40, 1e-40, 3e-40, 3e-40],
        [-5e-324, -5e-324, 5e-324, 5e-324],
        [3.5e38, 0.0, 1e39, 1.0],
        [-1e39, -1.0, -3.5e38, 0.0],
        [0.0, 0.0, 0.0, 0.0],
        [0.25, 0.25, 0.25, 0.25],
    ])
    b = np.concatenate([b, specials])
    q = np.concatenate([boxes64(rng, 500, 0.1), specials, specials + 1e-45,
                        np.array([[2e-40, 2e-40, 2e-40, 2e-40], [-1e-320, -1.0, 0.0, 0.0],
                                  [3.4e38, 0.5, 3.6e38, 0.5], [-np.inf, -np.inf, np.inf, np.inf]])])
    pts = np.concatenate([rng.random((300, 2)), specials[:, :2], specials[:, 2:],
                          np.array([[2e-40, 2e-40], [4e-324, 0.0], [5e38, 0.5], [0.25, 0.25]])])
    for mode in (THREAD, WAVE):
        ix = make_index(b, traversal=mode, leaf_size=4)
        for pred in (P.kIntersects, P.kContains, P.kWithin):
            got, n = collect(ix, pred, q)
            assert np.array_equal(got, oracle.bf_pairs64(int(pred), b, q)), (int(pred), mode)
        got, n = collect(ix, P.kContains, pts)
        assert np.array_equal(got, oracle.bf_pairs64(oracle.CONTAINS, b, pts)), mode


@pytest.mark.parametrize("policy", [N.RTS_REBUILD_NEVER, N.RTS_REBUILD_SAH, N.RTS_REBUILD_ALWAYS])
def test_f64_update_delete_match_oracle(oracle, data, policy):
    b, q, pts = data
    log(f"f64 update/delete policy={policy}")
    rng = np.random.default_rng(65 + policy)
    ix = make_index(b, batches=5, rebuild_policy=policy)
    ix.Build()
    ids = rng.choice(b.shape[0], 3000, replace=False).astype(np.int64)
    new = boxes64(rng, 3000, 0.02)
    ix.Update((dev(ids), dev(new)))
    dead = rng.choice(b.shape[0], 2000, replace=False).astype(np.int64)
    ix.Delete(dev(dead))
    ix.Sync()
    b2 = b.copy()
    b2[ids] = new
    alive = np.ones(b.shape[0], np.uint8)
    alive[dead] = 0
    for pred in (P.kIntersects, P.kWithin):
        got, n = collect(ix, pred, q)
        assert np.array_equal(got, oracle.bf_pairs64(int(pred), b2, q, alive=alive)), int(pred)
    got, n = collect(ix, P.kContains, pts)
    assert np.array_equal(got, oracle.bf_pairs64(oracle.CONTAINS, b2, pts, alive=alive))
    # an updated-then-deleted id stays deleted; a deleted id that is updated lives again
    revive = dead[:10]
    ix.Update((dev(revive), dev(b[revive])))
    alive[revive] = 1
    b2[revive] = b[revive]
    got, n = collect(ix, P.kIntersects, q)
    assert np.array_equal(got, oracle.bf_pairs64(oracle.INTERSECTS, b2, q, alive=alive))


def test_f64_autotune_keeps_results_exact(oracle, data):
    b, q, _ = data
    log("f64 autotune")
    ix = make_index(b)
    dq = dev(q)
    r = ix.Autotune(P.kIntersects, dq, sample=1024, reps=1)
    assert r["sample"] == 1024 and r["sample_results"] == oracle.bf_count64(
        oracle.INTERSECTS, b, q[:: q.shape[0] // 1024][:1024])
    got, n = collect(ix, P.kIntersects, dq)
    assert np.array_equal(got, oracle.bf_pairs64(oracle.INTERSECTS, b, q))
    rr = make_index(b, coord=N.RTS_COORD_F64_AS_F32)
    r2 = rr.Autotune(P.kIntersects, dq, sample=1024, reps=1)
    assert r2["sample_results"] == oracle.bf_count(
        oracle.INTERSECTS, b.astype(np.float32), q[:: q.shape[0] // 1024][:1024].astype(np.float32))


def test_coordinate_type_mismatch_is_rejected():
    log("f64 mismatched calls")
    L = N.lib()
    ix = R.SpatialIndex(torch.float64)
    ix.Init(R.Config())
    h = ix._handle()
    f = dev(np.zeros((4, 4), np.float32))
    d = dev(np.zeros((4, 4), np.float64))
    cnt = R.SharedValue()
    assert L.rts_index_insert(h, f.data_ptr(), 4, None) == 1
    assert L.rts_index_insert_d(h, d.data_ptr(), 4, None) == 0
    assert L.rts_index_query_envelopes(h, 0, f.data_ptr(), 4, 0, cnt.data(), None) == 1
    assert L.rts_index_update(h, None, 1, None) == 1
    fl = R.SpatialIndex()
    fl.Init(R.Config())
    assert L.rts_index_insert_d(fl._handle(), d.data_ptr(), 4, None) == 1
    assert L.rts_index_query_points_d(fl._handle(), 1, d.data_ptr(), 2, 0, cnt.data(), None) == 1
    c = N.rts_config()
    L.rts_config_default(ctypes.byref(c))
    c.coord = 7
    hh = ctypes.c_void_p()
    assert L.rts_index_create(ctypes.byref(c), ctypes.byref(hh)) == 1
    # served requests take the index's coordinate type
    from rtspatial_amd import serve as S
    with pytest.raises(ValueError, match="coordinate type"):
        S.run(ix, [S.Request(queries=dev(np.zeros((2, 4), np.float32)), predicate=P.kIntersects)])
    torch.cuda.synchronize()
