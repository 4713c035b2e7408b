"""GPU parity of line-segment queries (SpatialIndex::Query(p,
ArrayView<line_t>, ...), spatial_index.cuh:517-547) through the C ABI
(rts_index_query_lines / _lines_d), against the oracle's exact join.

The reference declares the query, but its intersection program reports no
results (shaders_intersects_line_query_2d.cu:14-41) and its only test has no
assertion (test/test_line_queries.h:9-29); README.md:59 lists line queries as
not yet implemented. Here kIntersects reports every data box the closed
segment meets, decided exactly; the oracle's segment predicate is pinned by
an exact rational restatement and an exact Liang-Barsky clip. kContains
reports nothing, as the reference's."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():
    pytest.skip("no GPU", allow_module_level=True)

import rtspatial_amd as R  # noqa: E402
from gpu_util import canon, dev, log, make_index  # noqa: E402

N = R._native
P = R.Predicate
THREAD, WAVE = N.RTS_TRAVERSAL_THREAD, N.RTS_TRAVERSAL_WAVE


def line_count(ix, pred, segs):
    c = R.SharedValue()
    c.set(0)
    ix.QueryLines(pred, dev(segs), c)
    return c.get()


def line_pairs(ix, pred, segs, capacity=None):
    s = dev(segs)
    if capacity is None:
        capacity = line_count(ix, pred, segs) + 1
    q = R.Queue(capacity)
    ix.QueryLines(pred, s, q)
    return canon(q.pairs()), q.size()


def check_pairs(ix, segs, want, what):
    n = line_count(ix, P.kIntersects, segs)
    assert n == want.shape[0], (what, n, want.shape[0])
    got, size = line_pairs(ix, P.kIntersects, segs)
    assert size == want.shape[0], what
    np.testing.assert_array_equal(got, want, err_msg=what)


def segments(rng, n, length, dtype=np.float32):
    a = rng.random((n, 2))
    b = a + (rng.random((n, 2)) - 0.5) * 2 * length
    return np.concatenate([a, b], axis=1).astype(dtype)


def edge_segments(rng, boxes, n, dtype=np.float32):
    """Segments through box corners, along box edges, degenerate, axis-aligned."""
    src = boxes[rng.integers(0, boxes.shape[0], n)].astype(dtype)
    out = np.empty((n, 4), dtype)
    kind = rng.integers(0, 5, n)
    for i in range(n):
        x0, y0, x1, y1 = src[i]
        k = kind[i]
        if k == 0:    # the box's diagonal, extended
            out[i] = [x0, y0, x1 + (x1 - x0), y1 + (y1 - y0)]
        elif k == 1:  # along the left edge
            out[i] = [x0, y0 - dtype(0.01), x0, y1 + dtype(0.01)]
        elif k == 2:  # along the top edge, from a corner outward
            out[i] = [x1, y1, x1 + dtype(0.02), y1]
        elif k == 3:  # a point on a corner
            out[i] = [x1, y0, x1, y0]
        else:         # anti-diagonal through a corner
            out[i] = [x0 - dtype(0.01), y0 + dtype(0.01), x0 + dtype(0.01), y0 - dtype(0.01)]
    return out


@pytest.fixture(scope="module")
def ref_data(oracle):
    # test/test_line_queries.h:9-29: 100K GenerateUniformBoxes(n, 0.5, 0.5)
    # boxes; the lines of GenerateUniformLines(n, w, h) (test_commons.h:69-97)
    # are the diagonals of GenerateUniformBoxes(n, w, h), here w = h = 0.1
    # as the envelope tests' queries (the reference's call omits them).
    boxes = oracle.ref_uniform_boxes(100_000, 0.5, 0.5)
    lines = oracle.ref_uniform_boxes(1000, 0.1, 0.1)
    return boxes, lines


@pytest.mark.parametrize("trav", [THREAD, WAVE])
def test_lines_reference_workload(oracle, ref_data, trav):
    log(f"lines reference workload traversal={trav}")
    boxes, lines = ref_data
    want = oracle.bf_pairs(oracle.INTERSECTS, boxes, lines, lines=True)
    assert want.shape[0] > 1000
    ix = make_index(boxes, traversal=trav)
    check_pairs(ix, lines, want, f"reference workload trav={trav}")
    assert line_count(ix, P.kContains, lines) == 0


@pytest.mark.parametrize("trav", [THREAD, WAVE])
@pytest.mark.parametrize("sort", [1, 2])
def test_lines_grid_exact_zero(oracle, trav, sort):
    # Boxes and segments on a 1/64 grid: many box corners lie exactly on a
    # segment's line (orientation exactly 0, the exact path decides).
    log(f"lines grid traversal={trav} sort={sort}")
    rng = np.random.default_rng(900)
    lo = rng.integers(0, 60, (20_000, 2))
    hi = lo + rng.integers(0, 4, (20_000, 2))
    boxes = (np.concatenate([lo, hi], axis=1) / 64).astype(np.float32)
    segs = (rng.integers(0, 64, (3000, 4)) / 64).astype(np.float32)
    segs[:500, 2:] = segs[:500, :2] + (rng.integers(-8, 9, (500, 2)) / 64)
    want = oracle.bf_pairs(oracle.INTERSECTS, boxes, segs, lines=True)
    ix = make_index(boxes, traversal=trav, sort_queries=sort)
    check_pairs(ix, segs, want, f"grid trav={trav} sort={sort}")


@pytest.mark.parametrize("trav", [THREAD, WAVE])
def test_lines_edges_and_degenerate(oracle, trav):
    log(f"lines edge cases traversal={trav}")
    rng = np.random.default_rng(901)
    boxes = oracle.ref_uniform_boxes(30_000, 0.02, 0.02)
    segs = np.concatenate([edge_segments(rng, boxes, 3000), segments(rng, 1000, 0.05)])
    segs[5] = [np.nan, 0.5, 0.6, 0.6]  # meets nothing
    want = oracle.bf_pairs(oracle.INTERSECTS, boxes, segs, lines=True)
    assert not np.any(want[:, 0] == 5)
    ix = make_index(boxes, traversal=trav)
    check_pairs(ix, segs, want, f"edges trav={trav}")


@pytest.mark.parametrize("trav", [THREAD, WAVE])
def test_lines_f64_exact(oracle, trav):
    # Double index, exact: segment ends and box corners within 1e-13 of one
    # another, far below a float ulp, so the float bounds cannot decide.
    log(f"lines f64 traversal={trav}")
    rng = np.random.default_rng(902)
    lo = rng.random((30_000, 2))
    boxes = np.concatenate([lo, lo + rng.random((30_000, 2)) * 0.01], axis=1)
    segs = np.concatenate([edge_segments(rng, boxes, 2000, np.float64),
                           segments(rng, 1000, 0.05, np.float64)])
    segs[:2000] += rng.choice([-1e-13, 0.0, 1e-13], size=(2000, 4))
    want = oracle.bf_pairs64(oracle.INTERSECTS, boxes, segs, lines=True)
    ix = make_index(boxes, traversal=trav)
    check_pairs(ix, segs, want, f"f64 trav={trav}")
    # the float answer on rounded inputs differs: the double path is not a float one
    wantf = oracle.bf_pairs(oracle.INTERSECTS, boxes.astype(np.float32),
                            segs.astype(np.float32), lines=True)
    assert wantf.shape != want.shape or not np.array_equal(wantf, want)


def test_lines_f64_as_f32(oracle):
    # coord RTS_COORD_F64_AS_F32: coordinates rounded to nearest floats, the
    # float predicate on the rounded values.
    log("lines f64 rounded to f32")
    rng = np.random.default_rng(903)
    lo = rng.random((20_000, 2))
    boxes = np.concatenate([lo, lo + rng.random((20_000, 2)) * 0.01], axis=1)
    segs = segments(rng, 2000, 0.05, np.float64)
    want = oracle.bf_pairs(oracle.INTERSECTS, boxes.astype(np.float32),
                           segs.astype(np.float32), lines=True)
    ix = make_index(boxes, coord=N.RTS_COORD_F64_AS_F32)
    check_pairs(ix, segs, want, "f64 as f32")


def test_lines_mutations_and_delta_tree(oracle):
    # Deleted boxes never match; boxes inserted after the build live in the
    # small tree of recent inserts, which the segment walk also visits.
    log("lines after delete / update / delta insert")
    rng = np.random.default_rng(904)
    boxes = oracle.ref_uniform_boxes(50_000, 0.01, 0.01)
    extra = oracle.ref_uniform_boxes(500, 0.01, 0.01)
    segs = segments(rng, 2000, 0.05)
    ix = make_index(boxes)
    ix.Build()
    dead = rng.choice(50_000, 5000, replace=False).astype(np.uint64)
    ix.Delete(dev(dead.astype(np.int64)))
    ix.Insert(dev(extra))
    allb = np.concatenate([boxes, extra])
    alive = np.ones(allb.shape[0], np.uint8)
    alive[dead] = 0
    want = oracle.bf_pairs(oracle.INTERSECTS, allb, segs, alive=alive, lines=True)
    check_pairs(ix, segs, want, "delta + deletes")
    assert ix.stats()["delta_boxes"] == 500  # the queries walked the small tree too


def test_lines_argument_rules(oracle):
    log("lines argument rules")
    boxes = oracle.ref_uniform_boxes(1000, 0.1, 0.1)
    segs = oracle.ref_uniform_boxes(100, 0.1, 0.1)
    ix = make_index(boxes)
    assert line_count(ix, P.kContains, segs) == 0
    with pytest.raises(N.RtsError, match="RTS_INTERSECTS or RTS_CONTAINS"):
        line_count(ix, P.kWithin, segs)
    with pytest.raises(TypeError):
        ix.QueryLines(P.kIntersects, dev(segs.astype(np.float64)), R.SharedValue())
    empty = R.SpatialIndex()
    empty.Init(R.Config())
    assert line_count(empty, P.kIntersects, segs) == 0
    assert line_count(ix, P.kIntersects, segs[:0]) == 0


def test_lines_large_properties(oracle):
    # 1M boxes x 200K segments: both traversals, both query orders and the
    # collecting handler agree on the count; a sample of queries matches the
    # oracle pair for pair.
    log("lines 1M x 200K properties")
    rng = np.random.default_rng(905)
    boxes = oracle.ref_uniform_boxes(1_000_000, 0.002, 0.002)
    segs = segments(rng, 200_000, 0.01)
    counts = set()
    for trav in (THREAD, WAVE):
        for sort in (1, 2):
            ix = make_index(boxes, traversal=trav, sort_queries=sort)
            counts.add(line_count(ix, P.kIntersects, segs))
    assert len(counts) == 1
    total = counts.pop()
    got, size = line_pairs(ix, P.kIntersects, segs)
    assert size == total and got.shape[0] == total
    sample = np.sort(rng.choice(200_000, 300, replace=False))
    want = oracle.bf_pairs(oracle.INTERSECTS, boxes, segs[sample], lines=True)
    want[:, 0] = sample[want[:, 0]]
    mine = got[np.isin(got[:, 0], sample)]
    np.testing.assert_array_equal(mine, want)
