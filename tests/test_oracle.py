"""CPU tests of the oracle: it must reproduce the reference tests' known
answers and agree with independent checks before it may check the GPU path."""
import json
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


@pytest.fixture(scope="module")
def ka():
    with open(os.path.join(GOLDEN, "known_answers.json")) as f:
        return json.load(f)


@pytest.fixture(scope="module")
def ref100k(oracle):
    return oracle.ref_uniform_boxes(100000, 0.5, 0.5)


# ---------------------------------------------------------------- generators --
def test_generator_streams_are_libstdcxx_mt19937(oracle):
    # Shape and range of the std::mt19937(n) streams of test/test_commons.h:22-67;
    # that they are the reference's exact streams is pinned by the known
    # answers below, which no other stream reproduces.
    p = oracle.ref_uniform_points(3)
    assert p.dtype == np.float32 and p.shape == (3, 2)
    assert np.all((p >= 0) & (p < 1))
    b = oracle.ref_uniform_boxes(1000, 0.1, 0.1)
    assert np.all(b[:, 2] >= b[:, 0]) and np.all(b[:, 3] >= b[:, 1])
    assert np.all(b[:, 2] <= 1.0) and np.all(b[:, 2] - b[:, 0] <= np.float32(0.1) + 1e-7)


def test_update_ids_distinct_and_deterministic(oracle, ka):
    ids = oracle.ref_update_ids(100000, 5000)
    assert len(set(ids.tolist())) == 5000 and ids.max() < 100000
    assert ids[:5].tolist() == ka["intersects_after_update_5pct"]["update_ids_first5"]


# ------------------------------------------------------------- known answers --
def test_point_contains_known_answer(oracle, ref100k):
    # test/test_point_queries.h:9-28,30-60 assert 4224111.
    pts = oracle.ref_uniform_points(1000)
    assert oracle.bf_count(oracle.CONTAINS, ref100k, pts) == 4224111


def test_intersects_known_answer_and_reference_extra_pair(oracle, ref100k):
    # test/test_envelope_queries.h:38-58 asserts 6549771. The reference's ray
    # formulation (ray_params.h:137-166) reproduces it exactly; the exact
    # closed predicate gives one pair fewer: (query 362, box 76881), where the
    # query's min y is 1 ulp above the box's max y.
    q = oracle.ref_uniform_boxes(1000, 0.1, 0.1)
    exact = oracle.bf_count(oracle.INTERSECTS, ref100k, q)
    ref, extra, missing = oracle.refray_intersects(ref100k, q)
    assert ref == 6549771
    assert exact == 6549770
    assert extra.tolist() == [[362, 76881]] and missing.size == 0
    g, qq = ref100k[76881], q[362]
    assert qq[1] == np.nextafter(g[3], np.float32(2)), "query min y is one ulp above box max y"
    assert not oracle.evaluate(oracle.INTERSECTS, g, qq)


def test_tiny_point_case(oracle, ka):
    c = ka["tiny_points"]
    boxes = np.array(c["boxes"], np.float32)
    pts = np.array(c["points"], np.float32)
    assert oracle.bf_count(oracle.CONTAINS, boxes, pts) == c["reference"] == 5
    assert oracle.bf_pairs(oracle.CONTAINS, boxes, pts).tolist() == c["pairs_query_box"]


def test_tiny_intersects_and_delete(oracle, ka):
    c = ka["tiny_intersects_delete"]
    boxes = np.array(c["boxes"], np.float32)
    q = np.array(c["queries"], np.float32)
    assert oracle.bf_count(oracle.INTERSECTS, boxes, q) == 2
    alive = np.array([1, 0], np.uint8)
    assert oracle.bf_count(oracle.INTERSECTS, boxes, q, alive=alive) == 1


def test_update_count_matches_rebuilt_array(oracle, ref100k, ka):
    # test/test_envelope_queries.h:93-164: the updated index must answer like a
    # fresh index over the updated array; the oracle evaluates the latter.
    ids = oracle.ref_update_ids(100000, 5000).astype(np.int64)
    new = oracle.ref_uniform_boxes(5000, 0.5, 0.5)
    b = ref100k.copy()
    b[ids] = new
    q = oracle.ref_uniform_boxes(1000, 0.1, 0.1)
    assert oracle.bf_count(oracle.INTERSECTS, b, q) == ka["intersects_after_update_5pct"]["exact"]


# ----------------------------------------------------- closed-form predicates --
@pytest.mark.parametrize("box,pt,want", [
    ([0, 0, 1, 1], [0, 0], True),          # corner: closed
    ([0, 0, 1, 1], [1, 1], True),
    ([0, 0, 1, 1], [1, 0.5], True),        # edge
    ([0, 0, 1, 1], [np.nextafter(np.float32(1), np.float32(2)), 0.5], False),
    ([0.5, 0.5, 0.5, 0.5], [0.5, 0.5], True),   # degenerate box = point
    ([np.nan] * 4, [0.5, 0.5], False),           # deleted box
])
def test_point_predicate_closed_form(oracle, box, pt, want):
    assert oracle.evaluate(oracle.CONTAINS, np.float32(box), np.float32(pt)) == want


@pytest.mark.parametrize("g,q,inter,cont,within", [
    ([0, 0, 1, 1], [1, 1, 2, 2], True, False, False),       # touching corners
    ([0, 0, 1, 1], [0.2, 0.2, 0.4, 0.4], True, True, False),
    ([0.2, 0.2, 0.4, 0.4], [0, 0, 1, 1], True, False, True),
    ([0, 0, 1, 1], [0, 0, 1, 1], True, True, True),          # equal boxes
    ([0, 0, 1, 1], [1.5, 0, 2, 1], False, False, False),
    ([0, 0, 1, 1], [0, 1.0000001, 1, 2], False, False, False),
    ([0, 0, 0, 1], [0, 0, 0, 1], True, True, True),          # zero width
    ([np.nan] * 4, [0, 0, 1, 1], False, False, False),
])
def test_envelope_predicates_closed_form(oracle, g, q, inter, cont, within):
    g, q = np.float32(g), np.float32(q)
    assert oracle.evaluate(oracle.INTERSECTS, g, q) == inter
    assert oracle.evaluate(oracle.CONTAINS, g, q) == cont
    assert oracle.evaluate(oracle.WITHIN, g, q) == within


def test_predicates_match_numpy_restatement(oracle):
    # A second, vectorised statement of the closed forms on random inputs
    # with many ties (coordinates on a coarse grid).
    rng = np.random.default_rng(7)
    g = np.sort(rng.integers(0, 8, (400, 2, 2)), axis=1).astype(np.float32) / 8
    q = np.sort(rng.integers(0, 8, (50, 2, 2)), axis=1).astype(np.float32) / 8
    g = g.transpose(0, 2, 1).reshape(-1, 4)[:, [0, 2, 1, 3]]
    q = q.transpose(0, 2, 1).reshape(-1, 4)[:, [0, 2, 1, 3]]
    G, Q = g[None, :, :], q[:, None, :]
    inter = (Q[..., 0] <= G[..., 2]) & (Q[..., 2] >= G[..., 0]) & (Q[..., 1] <= G[..., 3]) & (Q[..., 3] >= G[..., 1])
    cont = (Q[..., 0] >= G[..., 0]) & (Q[..., 2] <= G[..., 2]) & (Q[..., 1] >= G[..., 1]) & (Q[..., 3] <= G[..., 3])
    within = (G[..., 0] >= Q[..., 0]) & (G[..., 2] <= Q[..., 2]) & (G[..., 1] >= Q[..., 1]) & (G[..., 3] <= Q[..., 3])
    for pred, m in ((oracle.INTERSECTS, inter), (oracle.CONTAINS, cont), (oracle.WITHIN, within)):
        want = np.argwhere(m).astype(np.uint32)
        assert np.array_equal(oracle.bf_pairs(pred, g, q), want)
    p = (rng.integers(0, 9, (60, 2)) / 8).astype(np.float32)
    P = p[:, None, :]
    pin = (P[..., 0] >= G[..., 0]) & (P[..., 0] <= G[..., 2]) & (P[..., 1] >= G[..., 1]) & (P[..., 1] <= G[..., 3])
    assert np.array_equal(oracle.bf_pairs(oracle.CONTAINS, g, p), np.argwhere(pin).astype(np.uint32))


def test_planted_strict_inequality_is_caught(oracle):
    # A plausible mistake -- an open interval instead of a closed one -- must
    # change the answer on the tiny case the reference asserts.
    boxes = np.array([[0, 0, 1, 1], [0.5, -0.5, 1.5, 0.5]], np.float32)
    q = np.array([[0.75, 0.25, 1.25, 0.75]], np.float32)
    touching = np.array([[1, 0.5, 2, 2]], np.float32)
    assert oracle.bf_count(oracle.INTERSECTS, boxes, q) == 2
    assert oracle.bf_count(oracle.INTERSECTS, boxes, touching) == 2   # closed
    strict = sum(bool(t[0] < b[2] and t[2] > b[0] and t[1] < b[3] and t[3] > b[1])
                 for b in boxes for t in touching)
    assert strict != 2


# ------------------------------------------------------------------- R-tree --
@pytest.mark.parametrize("pred", [0, 1, 2])
def test_rtree_matches_brute_force(oracle, pred):
    rng = np.random.default_rng(11 + pred)
    b = np.concatenate([
        oracle.ref_uniform_boxes(20000, 0.05, 0.05),
        np.repeat(np.float32([[0.3, 0.3, 0.31, 0.31]]), 500, axis=0),   # duplicates
        np.float32([[0.5, 0.5, 0.5, 0.5], [0, 0, 1, 1]]),
    ])
    alive = (rng.random(b.shape[0]) > 0.1).astype(np.uint8)
    q = oracle.ref_uniform_boxes(500, 0.1, 0.1)
    t = oracle.RTree(b, alive=alive)
    assert np.array_equal(t.pairs(pred, q), oracle.bf_pairs(pred, b, q, alive=alive))
    pts = oracle.ref_uniform_points(700)
    assert np.array_equal(t.pairs(oracle.CONTAINS, pts), oracle.bf_pairs(oracle.CONTAINS, b, pts, alive=alive))


def test_rtree_handles_inverted_boxes(oracle):
    rng = np.random.default_rng(5)
    b = oracle.ref_uniform_boxes(5000, 0.05, 0.05)
    flip = rng.random(5000) < 0.4
    b[flip] = b[flip][:, [2, 3, 0, 1]]
    q = oracle.ref_uniform_boxes(300, 0.2, 0.2)
    t = oracle.RTree(b)
    for pred in (0, 1, 2):
        assert np.array_equal(t.pairs(pred, q), oracle.bf_pairs(pred, b, q))


def test_rtree_known_answers(oracle, ref100k):
    t = oracle.RTree(ref100k)
    assert t.count(oracle.CONTAINS, oracle.ref_uniform_points(1000)) == 4224111
    assert t.count(oracle.INTERSECTS, oracle.ref_uniform_boxes(1000, 0.1, 0.1)) == 6549770


def test_rtree_fanout_and_empty(oracle):
    b = oracle.ref_uniform_boxes(3000, 0.05, 0.05)
    q = oracle.ref_uniform_boxes(100, 0.1, 0.1)
    want = oracle.bf_pairs(oracle.INTERSECTS, b, q)
    for fan in (2, 4, 16, 64):
        assert np.array_equal(oracle.RTree(b, fanout=fan).pairs(oracle.INTERSECTS, q), want)
    empty = oracle.RTree(b, alive=np.zeros(3000, np.uint8))
    assert empty.count(oracle.INTERSECTS, q) == 0


# --------------------------------------------------------------- summaries --
def test_splitmix64_reference_value(oracle):
    # First output of SplitMix64 (Steele, Lea, Flood 2014) from state 0.
    assert oracle.splitmix64(0) == 0xE220A8397B1DCDAF
    assert oracle._splitmix_np(np.array([0], np.uint64))[0] == 0xE220A8397B1DCDAF


def test_summaries_agree(oracle):
    b = oracle.ref_uniform_boxes(5000, 0.05, 0.05)
    q = oracle.ref_uniform_boxes(300, 0.1, 0.1)
    total, cnt, hs, hx = oracle.bf_summary(oracle.INTERSECTS, b, q)
    pairs = oracle.bf_pairs(oracle.INTERSECTS, b, q)
    c2, s2, x2 = oracle.summarize_pairs(pairs, q.shape[0])
    assert total == pairs.shape[0]
    assert np.array_equal(cnt, c2) and np.array_equal(hs, s2) and np.array_equal(hx, x2)
    t2, cnt3, hs3, hx3 = oracle.RTree(b).summary(oracle.INTERSECTS, q)
    assert t2 == total and np.array_equal(cnt3, cnt) and np.array_equal(hs3, hs) and np.array_equal(hx3, hx)


def test_factored_checksum_functions(oracle):
    # ghash(b) = splitmix64(b ^ salt): at b = salt it is splitmix64(0), the
    # published first SplitMix64 output; the C and numpy forms agree, and
    # qmult is odd.
    salt = 0xD1B54A32D192ED03
    assert oracle.lib().orc_ghash(salt) == 0xE220A8397B1DCDAF
    vals = np.array([0, 1, 2, 3, 1 << 31, (1 << 32) - 1, 12345678, salt & 0xFFFFFFFF], np.uint64)
    for v in vals:
        assert oracle.lib().orc_ghash(int(v)) == int(oracle.ghash_np(v))
        assert oracle.lib().orc_qmult(int(v)) == int(oracle.qmult_np(v))
        assert int(oracle.qmult_np(v)) & 1
    # A request's checksum is the per-pair sum of qmult(q) * ghash(b).
    rng = np.random.default_rng(3)
    pairs = np.stack([rng.integers(0, 50, 4000), rng.integers(0, 1 << 31, 4000)], 1)
    cnt, gs = oracle.pair_checksums(pairs, 50)
    with np.errstate(over="ignore"):
        direct = int(np.sum(oracle.qmult_np(pairs[:, 0]) * oracle.ghash_np(pairs[:, 1]),
                            dtype=np.uint64))
    assert oracle.request_checksum(gs) == direct
    assert int(cnt.sum()) == 4000
    # One substituted box id changes the checksum.
    p2 = pairs.copy()
    p2[17, 1] += 1
    assert oracle.request_checksum(oracle.pair_checksums(p2, 50)[1]) != direct


def _adversarial_boxes(rng, n):
    """Boxes on a 1/64 grid (edges that exactly touch grid-aligned queries),
    some zero-width, some with swapped corners, some with a NaN coordinate."""
    g = rng.integers(0, 64, size=(n, 2)).astype(np.float32) / np.float32(64)
    w = rng.integers(0, 4, size=(n, 2)).astype(np.float32) / np.float32(64)
    b = np.concatenate([g, g + w], 1).astype(np.float32)
    sw = rng.random(n) < 0.05
    b[sw] = b[sw][:, [2, 3, 0, 1]]
    nan = rng.random(n) < 0.03
    b[nan, rng.integers(0, 4, size=int(nan.sum()))] = np.nan
    return b


def test_rtree_checksums_equal_brute_force(oracle):
    # Whole-subtree counting must agree with enumerating every pair: grid
    # queries touch box edges exactly (closed intervals), and NaN, swapped,
    # zero-width and dead boxes are all present; several fan-outs move the
    # shortcut to different levels.
    rng = np.random.default_rng(11)
    b = _adversarial_boxes(rng, 6000)
    alive = (rng.random(6000) > 0.1).astype(np.uint8)
    qg = rng.integers(0, 64, size=(400, 2)).astype(np.float32) / np.float32(64)
    qw = rng.integers(0, 24, size=(400, 2)).astype(np.float32) / np.float32(64)
    qe = np.concatenate([qg, qg + qw], 1).astype(np.float32)
    qe[:10] = np.float32([0, 0, 1, 1])           # covers every box: the root is inside
    qe[10:15] = qe[10:15][:, [2, 3, 0, 1]]       # swapped query corners
    pts = rng.integers(0, 65, size=(300, 2)).astype(np.float32) / np.float32(64)
    for fan in (2, 4, 16):
        tree = oracle.RTree(b, alive=alive, fanout=fan)
        for pred, q in ((oracle.INTERSECTS, qe), (oracle.WITHIN, qe), (oracle.CONTAINS, qe),
                        (oracle.CONTAINS, pts)):
            want = oracle.pair_checksums(oracle.bf_pairs(pred, b, q, alive=alive), q.shape[0])
            total, cnt, gs = tree.checksums(pred, q)
            assert np.array_equal(cnt, want[0]) and np.array_equal(gs, want[1]), (fan, pred)
            assert total == int(want[0].sum())


def test_rtree_update_equals_rebuild(oracle):
    # In-place updates (moves, NaN deletes, revivals) plus a refit answer
    # exactly as a tree built over the updated boxes.
    rng = np.random.default_rng(12)
    b = _adversarial_boxes(rng, 5000)
    alive = (rng.random(5000) > 0.1).astype(np.uint8)
    tree = oracle.RTree(b, alive=alive)
    cur = b.copy()
    for step in range(3):
        ids = rng.choice(5000, 700, replace=False).astype(np.uint64)
        nb = _adversarial_boxes(rng, 700)
        nb[:50] = np.nan
        applied = tree.update(ids, nb)
        live = alive[ids.astype(np.int64)] == 1
        assert applied == int(live.sum())
        cur[ids[live].astype(np.int64)] = nb[live]
        q = _adversarial_boxes(rng, 300)
        q = np.nan_to_num(q, nan=0.5)
        fresh = oracle.RTree(cur, alive=alive)
        for pred in (oracle.INTERSECTS, oracle.WITHIN, oracle.CONTAINS):
            got = tree.checksums(pred, q)
            want = fresh.checksums(pred, q)
            assert got[0] == want[0] and np.array_equal(got[1], want[1])
            assert np.array_equal(got[2], want[2])
            assert np.array_equal(tree.pairs(pred, q), oracle.bf_pairs(pred, cur, q, alive=alive))
    assert tree.delete(np.arange(5000, dtype=np.uint64)) == int(alive.sum())
    assert tree.count(oracle.INTERSECTS, np.float32([[0, 0, 1, 1]])) == 0


# ----------------------------------------------------------------- fixtures --
def test_wkt_fixtures_consistent(oracle):
    d = np.load(os.path.join(GOLDEN, "wkt_1000.npz"))
    for bk in ("uniform", "gaussian"):
        boxes = d[f"{bk}_boxes"]
        assert boxes.shape == (1000, 4) and np.all(boxes[:, 2] >= boxes[:, 0])
        for qk in ("uniform", "gaussian"):
            got = oracle.bf_pairs(oracle.CONTAINS, boxes, d[f"{qk}_points"])
            assert np.array_equal(got, d[f"pairs_{bk}_boxes_{qk}_points_contains"])
            for name, pred in (("intersects", 0), ("contains", 1), ("within", 2)):
                got = oracle.bf_pairs(pred, boxes, d[f"{qk}_boxes"])
                assert np.array_equal(got, d[f"pairs_{bk}_boxes_{qk}_boxes_{name}"])
    # within is contains with the roles of the two sets swapped
    a = d["pairs_uniform_boxes_gaussian_boxes_within"]
    b = d["pairs_gaussian_boxes_uniform_boxes_contains"]
    key = lambda p: np.sort(np.ascontiguousarray(p).view(np.uint64).ravel())
    assert np.array_equal(key(a[:, ::-1]), key(b))


# ------------------------------------------------------- double coordinates --
def test_f64_known_answer(oracle, ka):
    # test/test_point_queries.h:92-120 (fp64_contains_point) asserts 5 on the
    # tiny case's coordinates as doubles.
    c = ka["tiny_points"]
    boxes = np.array(c["boxes"], np.float64)
    pts = np.array(c["points"], np.float64)
    assert oracle.bf_count64(oracle.CONTAINS, boxes, pts) == 5
    assert oracle.bf_pairs64(oracle.CONTAINS, boxes, pts).tolist() == c["pairs_query_box"]


def test_f64_equals_f32_on_float_exact_inputs(oracle, ref100k):
    # Floats are exact doubles, so on float inputs the double join must equal
    # the (known-answer-pinned) float join, ties and deletions included.
    rng = np.random.default_rng(31)
    b = ref100k[:20000]
    alive = (rng.random(b.shape[0]) > 0.1).astype(np.uint8)
    q = oracle.ref_uniform_boxes(300, 0.1, 0.1)
    q[:20] = b[rng.integers(0, b.shape[0], 20)]  # exact equalities
    for pred in (oracle.INTERSECTS, oracle.CONTAINS, oracle.WITHIN):
        assert np.array_equal(oracle.bf_pairs64(pred, b.astype(np.float64), q.astype(np.float64),
                                                alive=alive),
                              oracle.bf_pairs(pred, b, q, alive=alive))
    p = oracle.ref_uniform_points(400)
    assert np.array_equal(oracle.bf_pairs64(oracle.CONTAINS, b.astype(np.float64), p.astype(np.float64)),
                          oracle.bf_pairs(oracle.CONTAINS, b, p))


# Cases whose answer flips when coordinates are rounded to float: the double
# oracle must give the exact answer (evaluating in float would give the other).
# Rounding to nearest is monotone and every predicate is a conjunction of <=
# comparisons, so rounding can only add pairs: every flip is exactly false.
F64_FLIPS = [
    # (pred, box, query)
    (1, [0.0, 0.0, 0.1, 0.1], [float(np.float32(0.1)), 0.05]),  # point just right of x1
    (1, [0.1, 0.0, 0.2, 1.0], [float(np.nextafter(0.1, 0)), 0.5]),
    (1, [0.0, 0.1, 1.0, 0.2], [0.5, float(np.nextafter(0.2, 1))]),  # just above y1
    (0, [0.0, 0.0, 0.3, 0.3], [float(np.nextafter(0.3, 1)), 0.0, 0.5, 0.5]),  # gap of 1 ulp
    (0, [0.0, 0.0, 0.5, 0.5], [0.2, -1.0, 0.3, float(np.nextafter(0.0, -1))]),
    (1, [0.0, 0.0, 1.0, 1.0], [0.5, 0.5, 1.0 + 1e-12, 0.75]),  # sticks out by 1e-12
    (1, [0.0, 0.2, 1.0, 1.0], [0.5, 0.2 - 1e-14, 0.6, 0.75]),
    (2, [0.1 - 1e-12, 0.2, 0.3, 0.4], [0.1, 0.1, 0.5, 0.5]),   # starts left of the query
    (2, [0.2, 0.2, 0.5 + 1e-12, 0.4], [0.1, 0.1, 0.5, 0.5]),   # ends right of it
]


@pytest.mark.parametrize("pred,box,query", F64_FLIPS)
def test_f64_rounding_flips(oracle, pred, box, query):
    b, q = np.float64(box), np.float64(query)
    assert not oracle.evaluate64(pred, b, q)
    # the same case rounded to float matches
    assert oracle.evaluate(pred, np.float32(b), np.float32(q))


def test_f64_matches_numpy_restatement(oracle):
    # Vectorised double restatement on random doubles with near-ties.
    rng = np.random.default_rng(9)
    g = np.sort(rng.random((600, 2, 2)), axis=1)
    g = g.transpose(0, 2, 1).reshape(-1, 4)[:, [0, 2, 1, 3]].copy()
    q = np.sort(rng.random((80, 2, 2)), axis=1)
    q = q.transpose(0, 2, 1).reshape(-1, 4)[:, [0, 2, 1, 3]].copy()
    q[:40] = g[:40] + rng.choice([-1e-13, 0.0, 1e-13], size=(40, 4))  # within an ulp of float
    G, Q = g[None, :, :], q[:, None, :]
    inter = (Q[..., 0] <= G[..., 2]) & (Q[..., 2] >= G[..., 0]) & (Q[..., 1] <= G[..., 3]) & (Q[..., 3] >= G[..., 1])
    cont = (Q[..., 0] >= G[..., 0]) & (Q[..., 2] <= G[..., 2]) & (Q[..., 1] >= G[..., 1]) & (Q[..., 3] <= G[..., 3])
    within = (G[..., 0] >= Q[..., 0]) & (G[..., 2] <= Q[..., 2]) & (G[..., 1] >= Q[..., 1]) & (G[..., 3] <= Q[..., 3])
    for pred, m in ((oracle.INTERSECTS, inter), (oracle.CONTAINS, cont), (oracle.WITHIN, within)):
        assert np.array_equal(oracle.bf_pairs64(pred, g, q), np.argwhere(m).astype(np.uint32))


# ------------------------------------------------------------ line segments --
def _liang_barsky(seg, box):
    # An independent statement: clip the segment's parameter range [0, 1] by
    # the four slabs in exact rationals; it meets the box iff a range remains.
    from fractions import Fraction as F
    x1, y1, x2, y2 = (F(float(v)) for v in seg)
    bx0, by0, bx1, by1 = (F(float(v)) for v in box)
    t0, t1 = F(0), F(1)
    for p, q in ((-(x2 - x1), x1 - bx0), (x2 - x1, bx1 - x1), (-(y2 - y1), y1 - by0),
                 (y2 - y1, by1 - y1)):
        if p == 0:
            if q < 0:
                return False
        else:
            r = q / p
            if p < 0:
                t0 = max(t0, r)
            else:
                t1 = min(t1, r)
            if t0 > t1:
                return False
    return True


@pytest.mark.parametrize("seg,box,want", [
    ([0, 0, 1, 1], [0.4, 0.4, 0.6, 0.6], True),     # through the middle
    ([0, 0, 1, 1], [0.6, 0.0, 0.8, 0.3], False),    # beside the diagonal
    ([0, 0, 1, 1], [1, 1, 2, 2], True),              # touches a corner
    ([0, 0, 1, 0], [0.5, 0, 0.6, 1], True),          # along an edge
    ([0, 0, 1, 0], [0.5, 1e-30, 0.6, 1], False),     # just above it
    ([0.5, 0.5, 0.5, 0.5], [0, 0, 1, 1], True),      # a point inside
    ([2, 2, 2, 2], [0, 0, 1, 1], False),             # a point outside
    ([0, 1, 1, 0], [0.5, 0.5, 0.7, 0.7], True),      # anti-diagonal through the corner
    ([0, 1, 1, 0], [np.nextafter(0.5, 1), np.nextafter(0.5, 1), 0.7, 0.7], False),
    ([0, 0, 1, 1], [np.nan] * 4, False),
])
def test_segment_closed_form(oracle, seg, box, want):
    assert oracle.segment_meets_box(seg, box) == want
    if not np.isnan(box[0]):
        assert _liang_barsky(seg, box) == want


def test_segment_predicate_matches_liang_barsky(oracle):
    # Random segments and boxes on a coarse grid (many touching and collinear
    # cases) and near-touching float cases: two exact methods agree.
    rng = np.random.default_rng(17)
    for _ in range(3000):
        s = rng.integers(0, 9, 4) / 8
        b = np.sort(rng.integers(0, 9, (2, 2)), axis=0).T.reshape(-1)[[0, 2, 1, 3]] / 8
        assert oracle.segment_meets_box(s, b) == _liang_barsky(s, b), (s, b)
    for _ in range(2000):
        s = rng.random(4).astype(np.float32)
        lo = rng.random(2).astype(np.float32)
        b = np.concatenate([lo, lo + rng.random(2).astype(np.float32) * np.float32(0.05)])
        if rng.random() < 0.5:  # put a corner on the segment's line, nudged by an ulp
            t = np.float32(rng.random())
            b[0] = s[0] + t * (s[2] - s[0])
            b[1] = np.nextafter(s[1] + t * (s[3] - s[1]), np.float32(rng.choice([-1, 1])))
            b[2], b[3] = max(b[0], b[2]), max(b[1], b[3])
        assert oracle.segment_meets_box(s, b) == _liang_barsky(s, b), (s, b)


def test_segment_pairs_prefilter(oracle):
    rng = np.random.default_rng(18)
    b = oracle.ref_uniform_boxes(2000, 0.05, 0.05)
    segs = rng.random((60, 4)).astype(np.float32)
    segs[:, 2:] = segs[:, :2] + (segs[:, 2:] - 0.5) * 0.2
    got = oracle.segment_pairs(b, segs)
    want = [(q, g) for q in range(segs.shape[0]) for g in range(b.shape[0])
            if _liang_barsky(segs[q], b[g])]
    assert got.tolist() == [list(p) for p in want]


def _adversarial_segments(rng, n, dtype):
    """Segments and boxes where corners sit on, or an ulp off, the segment's
    line, across wide exponent ranges."""
    out = []
    for _ in range(n):
        kind = rng.integers(0, 4)
        s = rng.random(4).astype(dtype)
        if kind == 1:  # tiny coordinates
            s = (s * dtype(10.0) ** -rng.integers(5, 30)).astype(dtype)
        t = dtype(rng.random())
        px = s[0] + t * (s[2] - s[0])
        py = s[1] + t * (s[3] - s[1])
        if kind == 2:  # a corner one ulp beside the line
            py = np.nextafter(py, dtype(rng.choice([-1.0, 1.0])))
        w = dtype(rng.random() * 0.1) * (dtype(1e-20) if kind == 1 else dtype(1))
        corner = rng.integers(0, 4)
        x0 = px - (w if corner in (1, 2) else 0)
        y0 = py - (w if corner in (2, 3) else 0)
        b = np.array([x0, y0, x0 + w, y0 + w], dtype)
        if kind == 3:  # degenerate segment / degenerate box
            s[2:] = s[:2] if rng.random() < 0.5 else s[2:]
            b[2:] = b[:2] if rng.random() < 0.5 else b[2:]
        out.append((s, b))
    return out


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_c_segment_predicate_matches_rational(oracle, dtype):
    # The C oracle's line predicate (double filter + integer arithmetic)
    # against the exact rational one.
    rng = np.random.default_rng(19 if dtype == np.float32 else 20)
    ev = oracle.evaluate if dtype == np.float32 else oracle.evaluate64
    cases = _adversarial_segments(rng, 3000, dtype)
    hits = 0
    for s, b in cases:
        want = oracle.segment_meets_box(s, b)
        assert ev(oracle.INTERSECTS, b, s, lines=True) == want, (s, b)
        assert not ev(oracle.CONTAINS, b, s, lines=True)
        hits += want
    assert 0.2 * len(cases) < hits < 0.95 * len(cases)  # both answers well represented


def test_c_segment_join_matches_rational(oracle):
    rng = np.random.default_rng(21)
    b = oracle.ref_uniform_boxes(3000, 0.05, 0.05)
    segs = rng.random((80, 4)).astype(np.float32)
    segs[:, 2:] = segs[:, :2] + (segs[:, 2:] - 0.5) * 0.3
    segs[0] = [0.5, 0.5, 0.5, 0.5]
    segs[1] = [np.nan, 0.1, 0.2, 0.3]
    segs[2] = b[7, [0, 1, 2, 3]]  # a box's diagonal
    segs[3] = [b[9, 0], 0.0, b[9, 0], 1.0]  # along a box's left edge
    alive = np.ones(3000, np.uint8)
    alive[::7] = 0
    got = oracle.bf_pairs(oracle.INTERSECTS, b, segs, alive=alive, lines=True)
    assert got.tolist() == oracle.segment_pairs(b, segs, alive).tolist()
    assert oracle.bf_count(oracle.CONTAINS, b, segs, lines=True) == 0
    got64 = oracle.bf_pairs64(oracle.INTERSECTS, b.astype(np.float64), segs.astype(np.float64),
                              alive=alive, lines=True)
    assert got64.tolist() == got.tolist()  # float-exact inputs: same answer in double


def test_c_segment_predicate_grid(oracle):
    # Coarse-grid segments and boxes: many corners exactly on the line
    # (orientation exactly 0) and segments along box edges.
    rng = np.random.default_rng(22)
    for _ in range(3000):
        s = (rng.integers(0, 9, 4) / 8).astype(np.float32)
        b = (np.sort(rng.integers(0, 9, (2, 2)), axis=0).T.reshape(-1)[[0, 2, 1, 3]] / 8)
        b = b.astype(np.float32)
        want = _liang_barsky(s, b)
        assert oracle.evaluate(oracle.INTERSECTS, b, s, lines=True) == want, (s, b)
        assert oracle.evaluate64(oracle.INTERSECTS, b, s, lines=True) == want, (s, b)
