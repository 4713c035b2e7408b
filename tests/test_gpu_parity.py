"""GPU parity tests: the HIP LBVH path, called through the C ABI, against the
oracle (exact brute force / R-tree) on the same seeded inputs, and against the
reference tests' known answers (test/test_point_queries.h,
test/test_envelope_queries.h). Integer/index results must match bit-exactly."""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():
    pytest.skip("no GPU", allow_module_level=True)

import rtspatial_amd as R  # noqa: E402
from gpu_util import canon, collect, count, dev, log, make_index, pinned  # noqa: E402

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
P = R.Predicate
THREAD, WAVE = R._native.RTS_TRAVERSAL_THREAD, R._native.RTS_TRAVERSAL_WAVE
MODES = [THREAD, WAVE]


@pytest.fixture(scope="module")
def ka():
    with open(os.path.join(GOLDEN, "known_answers.json")) as f:
        return json.load(f)


@pytest.fixture(scope="module")
def ref_inputs(oracle):
    return {
        "boxes": oracle.ref_uniform_boxes(100000, 0.5, 0.5),
        "points": oracle.ref_uniform_points(1000),
        "queries": oracle.ref_uniform_boxes(1000, 0.1, 0.1),
    }


# ------------------------------------------------------- reference known answers
def test_tiny_point_case_pinned_host(ka):
    # test/test_point_queries.h:62-90: pinned host inputs, count 5.
    log("tiny point case")
    c = ka["tiny_points"]
    boxes = pinned(np.array(c["boxes"], np.float32))
    pts = pinned(np.array(c["points"], np.float32))
    ix = R.SpatialIndex()
    ix.Init(R.Config())
    ix.Insert(boxes)
    cnt = R.SharedValue()
    cnt.set(0)
    ix.Query(P.kContains, pts, cnt)
    assert cnt.get() == 5
    pairs, n = collect(ix, P.kContains, dev(np.array(c["points"], np.float32)))
    assert n == 5 and pairs.tolist() == c["pairs_query_box"]


def test_tiny_intersects_then_delete(ka):
    # test/test_envelope_queries.h:11-36 and :166-245: 2, then 1 after Delete({1}).
    log("tiny intersects + delete")
    c = ka["tiny_intersects_delete"]
    for compact in (False, True):
        ix = R.SpatialIndex()
        ix.Init(R.Config(compact=compact))
        ix.Insert(pinned(np.array(c["boxes"], np.float32)))
        q = pinned(np.array(c["queries"], np.float32))
        assert count(ix, P.kIntersects, q) == 2
        ix.Delete(dev(np.array([1], np.int64)))
        torch.cuda.synchronize()
        assert count(ix, P.kIntersects, q) == 1
        pairs, _ = collect(ix, P.kIntersects, q)
        assert pairs.tolist() == [[0, 0]]


@pytest.mark.parametrize("batches", [1, 10])
def test_point_contains_known_answer(ref_inputs, batches):
    # test/test_point_queries.h:9-60: 4224111, single and 10-batch insert.
    log(f"point known answer, batches={batches}")
    ix = make_index(ref_inputs["boxes"], batches=batches)
    assert count(ix, P.kContains, ref_inputs["points"]) == 4224111


@pytest.mark.parametrize("batches", [1, 10])
def test_intersects_known_answer_exact(ref_inputs, ka, batches):
    # test/test_envelope_queries.h:38-91 asserts 6549771, which includes one
    # 1-ulp false positive of the reference's ray formulation; the exact set
    # has 6549770 pairs and excludes (query 362, box 76881).
    log(f"intersects known answer, batches={batches}")
    ix = make_index(ref_inputs["boxes"], batches=batches)
    assert count(ix, P.kIntersects, ref_inputs["queries"]) == 6549770
    pairs, n = collect(ix, P.kIntersects, ref_inputs["queries"])
    assert n == 6549770
    extra = ka["intersects_100k_1k"]["extra_pairs_query_box"][0]
    sel = pairs[pairs[:, 0] == extra[0]]
    assert extra[1] not in sel[:, 1]


def test_update_refit_equals_rebuild(ref_inputs, ka, oracle):
    # test/test_envelope_queries.h:93-164: update 5% of ids after a 10-batch
    # insert; the refitted index must answer like a freshly built one.
    log("update 5% refit == rebuild")
    ids = oracle.ref_update_ids(100000, 5000).astype(np.int64)
    new = oracle.ref_uniform_boxes(5000, 0.5, 0.5)
    b2 = ref_inputs["boxes"].copy()
    b2[ids] = new
    fresh = make_index(b2)
    want = count(fresh, P.kIntersects, ref_inputs["queries"])
    assert want == ka["intersects_after_update_5pct"]["exact"]
    for policy in (R._native.RTS_REBUILD_NEVER, R._native.RTS_REBUILD_SAH, R._native.RTS_REBUILD_ALWAYS):
        ix = make_index(ref_inputs["boxes"], batches=10, rebuild_policy=policy)
        ix.Build()
        ix.Update((dev(ids), dev(new)))
        ix.Sync()
        got, _ = collect(ix, P.kIntersects, ref_inputs["queries"])
        exp, _ = collect(fresh, P.kIntersects, ref_inputs["queries"])
        assert np.array_equal(got, exp), f"policy {policy}"


# -------------------------------------------------------------- pair-set parity
@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("leaf_size", [1, 4, 8, 32])
def test_c1_point_pairs_exact(oracle, leaf_size, mode):
    # Config 1 at full size: 100k boxes x 100k points, sorted pair set vs brute force.
    log(f"c1 point pairs leaf_size={leaf_size} traversal={mode}")
    from rtspatial_amd import workloads as W
    b, q = W.config1()
    want = oracle.bf_pairs(oracle.CONTAINS, b, q)
    ix = make_index(b, leaf_size=leaf_size, traversal=mode)
    got, n = collect(ix, P.kContains, q)
    assert n == want.shape[0]
    assert np.array_equal(got, want)


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("pred", [P.kIntersects, P.kContains, P.kWithin])
def test_c2_style_envelope_pairs_exact(oracle, pred, mode):
    # Config 2 shapes at 200k x 100k: every predicate, sorted pair set vs R-tree.
    log(f"c2-style envelope pairs pred={pred.name} traversal={mode}")
    from rtspatial_amd import workloads as W
    b, q = W.config2(200_000, 100_000)
    if pred == P.kContains:
        # data boxes that contain a query: make the queries small
        q = W.uniform_boxes(np.random.default_rng(5), 100_000, 1e-4)
    want = oracle.RTree(b).pairs(int(pred), q)
    ix = make_index(b, traversal=mode)
    got, n = collect(ix, pred, q)
    assert n == want.shape[0] and n > 0
    assert np.array_equal(got, want)


@pytest.mark.parametrize("leaf_size", [1, 8, 64])
def test_c3_style_skewed_intersects_exact(oracle, leaf_size):
    # Config 3's distributions at 1M boxes x 20k queries (wave traversal, the
    # mode AUTO picks): sorted pair set vs R-tree.
    log(f"c3-style skewed intersects leaf_size={leaf_size}")
    from rtspatial_amd import workloads as W
    b, q = W.config3(1_000_000, 20_000)
    want = oracle.RTree(b).pairs(oracle.INTERSECTS, q)
    ix = make_index(b, leaf_size=leaf_size)
    got, n = collect(ix, P.kIntersects, q)
    assert n == want.shape[0] and n > 1_000_000
    assert np.array_equal(got, want)


def test_wave_output_is_grouped_in_query_order_and_appends(oracle):
    # The two-pass wave path writes each query's pairs contiguously, queries
    # in input order; a second Query on the same queue appends after the first.
    log("wave output grouping + append")
    from rtspatial_amd import workloads as W
    b, q = W.config2(100_000, 30_000)
    ix = make_index(b, traversal=WAVE)
    n1 = count(ix, P.kIntersects, q)
    queue = R.Queue(2 * n1 + 10)
    ix.Query(P.kIntersects, dev(q), queue)
    ix.Query(P.kIntersects, dev(q[:1000]), queue)
    n1000 = oracle.bf_count(oracle.INTERSECTS, b, q[:1000])
    assert queue.size() == n1 + n1000
    raw = queue.pairs().cpu().numpy()
    first, second = raw[:n1], raw[n1:]
    assert np.all(np.diff(first[:, 1]) >= 0)
    assert np.all(np.diff(second[:, 1]) >= 0)
    assert np.array_equal(canon(torch.from_numpy(second)), oracle.bf_pairs(oracle.INTERSECTS, b, q[:1000]))


@pytest.mark.parametrize("name", ["uniform", "gaussian"])
def test_wkt_fixture_pairs(name):
    # The reference's sample datasets (fixtures in tests/golden/wkt_1000.npz).
    log(f"wkt fixtures {name}")
    d = np.load(os.path.join(GOLDEN, "wkt_1000.npz"))
    boxes = d[f"{name}_boxes"]
    ix = make_index(boxes, leaf_size=2)
    for qk in ("uniform", "gaussian"):
        got, _ = collect(ix, P.kContains, d[f"{qk}_points"])
        assert np.array_equal(got, d[f"pairs_{name}_boxes_{qk}_points_contains"])
        for pn, pred in (("intersects", P.kIntersects), ("contains", P.kContains), ("within", P.kWithin)):
            got, _ = collect(ix, pred, d[f"{qk}_boxes"])
            assert np.array_equal(got, d[f"pairs_{name}_boxes_{qk}_boxes_{pn}"]), (qk, pn)


# ------------------------------------------------------------------ edge cases
def test_empty_index_and_empty_queries():
    log("empty cases")
    ix = R.SpatialIndex()
    ix.Init(R.Config())
    q = dev(np.float32([[0, 0, 1, 1]]))
    assert count(ix, P.kIntersects, q) == 0
    ix.Insert(dev(np.float32([[0, 0, 1, 1]])))
    assert count(ix, P.kIntersects, dev(np.zeros((0, 4), np.float32))) == 0
    assert count(ix, P.kIntersects, q) == 1


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("n", [1, 2, 3, 63, 64, 65, 4095, 4096, 4097])
def test_small_sizes_and_tile_edges(oracle, n, mode):
    log(f"small size n={n} traversal={mode}")
    rng = np.random.default_rng(n)
    from rtspatial_amd import workloads as W
    b = W.uniform_boxes(rng, n, 0.2)
    q = W.uniform_boxes(rng, 257, 0.3)
    ix = make_index(b, leaf_size=1, traversal=mode)
    for pred in (P.kIntersects, P.kContains, P.kWithin):
        got, _ = collect(ix, pred, q)
        assert np.array_equal(got, oracle.bf_pairs(int(pred), b, q))


@pytest.mark.parametrize("mode", MODES)
def test_duplicates_degenerate_and_boundaries(oracle, mode):
    # Identical boxes (equal Morton codes), zero-width boxes, boxes touching
    # the query only on an edge or a corner, points on edges.
    log("duplicates / degenerate / boundaries")
    grid = (np.arange(0, 9, dtype=np.float32) / 8)
    xs, ys = np.meshgrid(grid[:-1], grid[:-1])
    cells = np.stack([xs.ravel(), ys.ravel(), xs.ravel() + np.float32(0.125), ys.ravel() + np.float32(0.125)], 1)
    dup = np.repeat(np.float32([[0.25, 0.25, 0.5, 0.5]]), 3000, axis=0)
    degen = np.float32([[0.5, 0.5, 0.5, 0.5], [0.25, 0, 0.25, 1], [0, 0.75, 1, 0.75]])
    b = np.concatenate([cells, dup, degen]).astype(np.float32)
    q = np.concatenate([cells, np.float32([[0.5, 0.5, 0.5, 0.5], [0.25, 0.25, 0.5, 0.5], [0, 0, 1, 1]])])
    pts = np.stack(np.meshgrid(grid, grid), -1).reshape(-1, 2).astype(np.float32)
    for leaf in (1, 8):
        ix = make_index(b, leaf_size=leaf, traversal=mode)
        for pred in (P.kIntersects, P.kContains, P.kWithin):
            got, _ = collect(ix, pred, q)
            assert np.array_equal(got, oracle.bf_pairs(int(pred), b, q)), (leaf, pred)
        got, _ = collect(ix, P.kContains, pts)
        assert np.array_equal(got, oracle.bf_pairs(oracle.CONTAINS, b, pts))


@pytest.mark.parametrize("mode", MODES)
def test_inverted_boxes_match_exact_predicates(oracle, mode):
    # Boxes whose min corner exceeds the max corner on one or both axes still
    # go through the exact closed predicates; bounds and the inside-subtree
    # shortcut must not change their results.
    log(f"inverted boxes traversal={mode}")
    from rtspatial_amd import workloads as W
    rng = np.random.default_rng(17)
    b = W.uniform_boxes(rng, 30_000, 0.02)
    flip = rng.random(b.shape[0]) < 0.3
    b[flip] = b[flip][:, [2, 1, 0, 3]]
    flip2 = rng.random(b.shape[0]) < 0.2
    b[flip2] = b[flip2][:, [0, 3, 2, 1]]
    q = np.concatenate([W.uniform_boxes(rng, 500, 0.3), np.float32([[0, 0, 1, 1]])])
    pts = W.uniform_points(rng, 2000)
    ix = make_index(b, leaf_size=4, traversal=mode)
    for pred in (P.kIntersects, P.kContains, P.kWithin):
        got, _ = collect(ix, pred, q)
        assert np.array_equal(got, oracle.bf_pairs(int(pred), b, q)), pred
    got, _ = collect(ix, P.kContains, pts)
    assert np.array_equal(got, oracle.bf_pairs(oracle.CONTAINS, b, pts))


@pytest.mark.parametrize("mode", MODES)
def test_delete_update_sequence_matches_oracle(oracle, mode):
    # Deletes (including out-of-range ids, which are ignored), updates,
    # re-updates of deleted ids, inserts after deletes.
    log(f"delete/update sequence traversal={mode}")
    from rtspatial_amd import workloads as W
    rng = np.random.default_rng(3)
    b = W.uniform_boxes(rng, 50_000, 0.01)
    q = W.uniform_boxes(rng, 2000, 0.05)
    alive = np.ones(b.shape[0], np.uint8)
    ix = make_index(b, traversal=mode)
    for step in range(4):
        dels = rng.choice(b.shape[0], 3000, replace=False).astype(np.int64)
        ix.Delete(dev(np.concatenate([dels, [b.shape[0] + 5, 1 << 40]]).astype(np.int64)))
        alive[dels] = 0
        ups = rng.choice(b.shape[0], 2000, replace=False).astype(np.int64)
        nb = W.uniform_boxes(rng, 2000, 0.02)
        ix.Update((dev(ups), dev(nb)))
        b[ups] = nb
        alive[ups] = 1
        if step == 2:
            extra = W.uniform_boxes(rng, 5000, 0.01)
            ix.Insert(dev(extra))
            b = np.concatenate([b, extra])
            alive = np.concatenate([alive, np.ones(5000, np.uint8)])
        qq = np.concatenate([q, np.float32([[-1, -1, 2, 2], [0.2, 0.2, 0.7, 0.7]])])
        for pred in (P.kIntersects, P.kWithin):
            got, _ = collect(ix, pred, qq)
            assert np.array_equal(got, oracle.bf_pairs(int(pred), b, qq, alive=alive)), (step, pred)
            assert count(ix, pred, qq) == got.shape[0]
    ix.Sync()


@pytest.mark.parametrize("mode", MODES)
def test_queue_overflow_reports_exact_count(oracle, mode):
    log(f"queue overflow traversal={mode}")
    from rtspatial_amd import workloads as W
    b, q = W.config2(100_000, 20_000)
    want = oracle.bf_count(oracle.WITHIN, b, q)
    ix = make_index(b, traversal=mode)
    queue = R.Queue(1000)
    ix.Query(P.kWithin, dev(q), queue)
    assert queue.size() == want and queue.overflowed()
    stored = canon(queue.pairs())
    full = oracle.bf_pairs(oracle.WITHIN, b, q)
    s = set(map(tuple, full.tolist()))
    assert stored.shape[0] == 1000 and all(tuple(p) in s for p in stored.tolist())


def test_query_sorting_modes_agree(oracle):
    log("query ordering modes")
    from rtspatial_amd import workloads as W
    b, q = W.config2(200_000, 100_000)
    want = oracle.RTree(b).pairs(oracle.INTERSECTS, q)
    for mode in (1, 2):
        ix = make_index(b, sort_queries=mode)
        got, _ = collect(ix, P.kIntersects, q)
        assert np.array_equal(got, want), mode


def test_point_predicate_other_than_contains_rejected():
    ix = make_index(np.float32([[0, 0, 1, 1]]))
    with pytest.raises(R.RtsError):
        ix.Query(P.kIntersects, dev(np.float32([[0.5, 0.5]])), R.SharedValue())


# ---------------------------------------------------------------------- sort
@pytest.mark.parametrize("n,bits", [(1, 30), (1000, 8), (4097, 16), (1_000_003, 30), (3_000_000, 32)])
def test_radix_sort_matches_stable_argsort(n, bits):
    log(f"sort n={n} bits={bits}")
    rng = np.random.default_rng(n)
    keys = rng.integers(0, 1 << bits, size=n, dtype=np.uint64).astype(np.uint32)
    if n > 10:
        keys[: n // 3] = keys[0]  # many equal keys: stability matters
    vals = rng.integers(0, 1 << 32, size=n, dtype=np.uint64).astype(np.uint32)
    k, v = R.sort_pairs(dev(keys.view(np.int32)), dev(vals.view(np.int32)), key_bits=bits)
    order = np.argsort(keys, kind="stable")
    assert np.array_equal(k.cpu().numpy().view(np.uint32), keys[order])
    assert np.array_equal(v.cpu().numpy().view(np.uint32), vals[order])
    k2, v2 = R.sort_pairs(dev(keys.view(np.int32)), None, key_bits=bits)
    assert np.array_equal(v2.cpu().numpy().view(np.uint32), order.astype(np.uint32))


@pytest.mark.parametrize("policy", ["refit", "sah", "rebuild"])
def test_mutable_rounds_match_oracle_every_round(oracle, policy):
    # The mutable-index workload at 1/10 scale: 1M uniform boxes (sides
    # U[0, 1e-3]); each round deletes 100k live ids, inserts 100k fresh boxes,
    # shifts 100k other live boxes by U[-1e-3, 1e-3], then answers 100k
    # uniform points. Every policy must give the oracle's pair set each round.
    from rtspatial_amd import workloads as W
    pol = {"refit": R._native.RTS_REBUILD_NEVER, "sah": R._native.RTS_REBUILD_SAH,
           "rebuild": R._native.RTS_REBUILD_ALWAYS}[policy]
    log(f"mutable rounds policy={policy}")
    m = W.MutationRounds(1_000_000, 100_000, 100_000)
    ix = make_index(m.initial, rebuild_policy=pol)
    for rnd in range(5):
        del_ids, ins, upd_ids, upd, pts = m.next_round()
        before = ix.stats()
        ix.Delete(dev(del_ids))
        if rnd < 3:
            ix.Insert(dev(ins))
        else:
            # Rounds without inserts: the tree is refitted (or rebuilt under
            # the rebuild policy) instead of rebuilt for new ids.
            m.alive[-ins.shape[0]:] = False
        ix.Update((dev(upd_ids), dev(upd)))
        got, n = collect(ix, P.kContains, pts)
        want = oracle.RTree(m.boxes, alive=m.alive.astype(np.uint8)).pairs(oracle.CONTAINS, pts)
        assert n == want.shape[0], (rnd, n, want.shape[0])
        assert np.array_equal(got, want), rnd
        after = ix.stats()
        if rnd < 3:
            # One rebuild per round, and no refit wasted before it.
            assert after["builds"] == before["builds"] + 1
            assert after["refits"] == before["refits"]
        elif policy == "refit":
            assert after["refits"] == before["refits"] + 1 and after["builds"] == before["builds"]
        else:
            # rebuild; and sah, as 100k updates are over 1/200 of the boxes
            assert after["builds"] == before["builds"] + 1
            assert after["refits"] == before["refits"]
    ix.Sync()


def test_sah_policy_refits_small_updates_and_rebuilds_large_ones(oracle):
    # Under RTS_REBUILD_SAH, updates covering less than 1/200 of the boxes
    # since the last build are refitted; once they reach it, the tree is rebuilt.
    rng = np.random.default_rng(31)
    boxes = oracle.ref_uniform_boxes(200_000, 0.01, 0.01)
    ix = make_index(boxes)
    ix.Build()
    pts = rng.random((20_000, 2)).astype(np.float32)
    cur = boxes.copy()
    for n_upd, rebuilt in ((100, False), (400, False), (600, True), (10, False)):
        log(f"sah policy: {n_upd} updates")
        ids = rng.choice(200_000, n_upd, replace=False)
        nb = cur[ids] + np.float32(0.002)
        cur[ids] = nb
        before = ix.stats()
        ix.Update((dev(ids.astype(np.int64)), dev(nb)))
        got, _ = collect(ix, P.kContains, pts)
        assert np.array_equal(got, oracle.RTree(cur).pairs(oracle.CONTAINS, pts))
        after = ix.stats()
        assert after["builds"] == before["builds"] + (1 if rebuilt else 0), n_upd
        assert after["refits"] == before["refits"] + (0 if rebuilt else 1), n_upd


@pytest.mark.parametrize("kind", ["points", "intersects", "within"])
def test_autotune_keeps_results_exact(oracle, kind):
    # Every plan the tuner times must give the same results; after tuning,
    # queries run under the stored plan and still equal the oracle's.
    from rtspatial_amd import workloads as W
    log(f"autotune {kind}")
    rng = np.random.default_rng(21)
    means = W.mixture_means(rng)
    b = W.mixture_boxes(rng, means, 200_000, np.log(1e-3))
    if kind == "points":
        q, pred = W.mixture_boxes(rng, means, 30_000, np.log(1e-3))[:, :2].copy(), P.kContains
    else:
        q = W.mixture_boxes(rng, means, 30_000, np.log(5e-3))
        pred = P.kIntersects if kind == "intersects" else P.kWithin
    ix = make_index(b)
    want = oracle.RTree(b).pairs(int(pred), q)
    dq = dev(q)
    r = ix.Autotune(pred, dq, sample=4096, reps=2)
    # A strided sample: 8 plans (4 leaf sizes x 2 traversals), 5 when a
    # traversal mode is dropped after its first plan, 4 when the sample's
    # queries average more than 256 results (one query per lane is then not
    # timed at all); then the winner in both query orders on the whole set
    # (30k queries: cheap to run), 2 more.
    lane_hopeless = r["sample_results"] > 256 * r["sample"]
    lo_plans = 4 if lane_hopeless else 5
    assert r["sample"] == 4096 and lo_plans + 2 <= r["plans"] <= 10
    assert r["sort_queries"] in (1, 2)
    if lane_hopeless:
        assert r["plans"] == 6 and r["traversal"] == WAVE
    assert r["leaf_size"] in (4, 8, 16, 32) and r["traversal"] in (THREAD, WAVE)
    assert 0 < r["best_ms"] <= r["default_ms"]
    sample_want = oracle.RTree(b).count(int(pred), q[:: q.shape[0] // 4096][:4096])
    assert r["sample_results"] == sample_want
    if kind == "intersects":
        # The whole set as the sample: the order is tuned too (16 plans, 9 or
        # more; 8 without the per-lane plans), and the stored plan includes it.
        r2 = ix.Autotune(pred, dq, sample=q.shape[0], reps=1)
        hopeless2 = r2["sample_results"] > 256 * r2["sample"]
        assert r2["sample"] == q.shape[0] and (8 if hopeless2 else 9) <= r2["plans"] <= 16
        assert r2["sort_queries"] in (1, 2) and r2["sample_results"] == want.shape[0]
    got, n = collect(ix, pred, dq)
    assert n == want.shape[0] and np.array_equal(got, want)
    assert count(ix, pred, dq) == want.shape[0]
    ix.Sync()
