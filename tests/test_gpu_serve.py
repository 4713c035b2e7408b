"""GPU tests of the served (continuous batching) and plain per-request paths:
every request's pair set must equal the oracle's answer for that request
alone, on the index state of the request's epoch, whatever else runs beside
it (mixed predicates, uneven sizes, overlapping batches, arena overflows and
re-launches, interleaved updates)."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
os.environ.setdefault("RTS_SERVE_TRACE", "1")  # host-loop steps go to stderr (shown on failure)

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():
    pytest.skip("no GPU", allow_module_level=True)

import rtspatial_amd as R  # noqa: E402
from rtspatial_amd import _native as N  # noqa: E402
from rtspatial_amd import serve as S  # noqa: E402
from rtspatial_amd import workloads as W  # noqa: E402
from gpu_util import canon, collect, dev, log, make_index  # noqa: E402

P = R.Predicate


def make_trace(rng, n_req, means, max_size, rate_hz=None, mix=(0.4, 0.3, 0.3)):
    """Requests of the served mix: points, within, intersects; sizes
    log-uniform in [1, max_size]."""
    reqs, host = [], []
    t = 0.0
    for i in range(n_req):
        kind = rng.choice(3, p=list(mix))
        n = int(np.rint(np.exp(rng.uniform(0, np.log(max_size)))))
        if kind == 0:
            q = W.mixture_boxes(rng, means, n, np.log(5e-3))[:, :2].copy()
            pred = P.kContains
        else:
            q = W.mixture_boxes(rng, means, n, np.log(5e-3))
            pred = P.kWithin if kind == 1 else P.kIntersects
        if rate_hz:
            t += rng.exponential(1.0 / rate_hz)
        host.append((q, pred))
        reqs.append(S.Request(queries=dev(q), predicate=pred, arrival_s=t, collect=0))
    return reqs, host


def expected(oracle, boxes, alive, q, pred):
    return oracle.RTree(boxes, alive=alive).pairs(int(pred), q)


@pytest.fixture(scope="module")
def scene():
    rng = np.random.default_rng(77)
    means = W.mixture_means(rng)
    boxes = W.mixture_boxes(rng, means, 300_000, np.log(1e-3))
    return rng, means, boxes


def _collect_caps(oracle, boxes, host, reqs):
    tree = oracle.RTree(boxes)
    want = []
    for (q, pred), r in zip(host, reqs):
        w = tree.pairs(int(pred), q)
        want.append(w)
        r.collect = w.shape[0] + 8
    return want


GRAPH_SKIP = ("hipGraph replay of the batch sequence did not complete on the MI355X; "
              "the cause is not established, so the path is off and this case stays off the GPU")


@pytest.mark.parametrize("mode,graphs", [
    pytest.param("batched", 1, marks=pytest.mark.skip(reason=GRAPH_SKIP)),
    ("batched", 0), ("plain", 0)])
def test_each_request_alone_equals_in_mix(oracle, scene, mode, graphs):
    rng, means, boxes = scene
    log(f"serve mix mode={mode} graphs={graphs}")
    reqs, host = make_trace(np.random.default_rng(1), 80, means, 3000)
    want = _collect_caps(oracle, boxes, host, reqs)
    ix = make_index(boxes)
    res, st = S.run(ix, reqs, mode=mode, use_graphs=graphs, slots=3)
    assert st["requests"] == 80 and st["batches"] >= 1
    for i, (r, w) in enumerate(zip(res, want)):
        assert r["count"] == w.shape[0], i
        assert not r["truncated"]
        assert np.array_equal(canon(r["pairs"]), w), i


def test_poisson_arrivals_overlapping_batches_and_hashes(oracle, scene):
    # Open-loop arrivals; batched and plain runs report the same per-request
    # counts and pair hashes, and both equal the oracle's.
    rng, means, boxes = scene
    log("serve poisson arrivals + hashes")
    reqs, host = make_trace(np.random.default_rng(2), 120, means, 20000, rate_hz=2000.0)
    ix = make_index(boxes)
    rb, sb = S.run(ix, reqs, mode="batched", hash_results=1, slots=3)
    rp, sp = S.run(ix, reqs, mode="plain", hash_results=1)
    tree = oracle.RTree(boxes)
    for i, (q, pred) in enumerate(host):
        total, cnt, gs = tree.checksums(int(pred), q)
        h = oracle.request_checksum(gs)
        assert rb[i]["count"] == rp[i]["count"] == total, i
        assert rb[i]["hash"] == rp[i]["hash"] == h, i
        assert rb[i]["done_s"] >= rb[i]["admit_s"] >= reqs[i].arrival_s - 1e-9
    assert sb["queries"] == sp["queries"]


@pytest.mark.parametrize("mode", ["batched", "plain"])
def test_arena_overflow_relaunch_is_exact(oracle, scene, mode):
    # A tiny arena forces overflows, re-launches and chunk splits.
    rng, means, boxes = scene
    log(f"serve arena overflow mode={mode}")
    reqs, host = make_trace(np.random.default_rng(3), 30, means, 400)
    want = _collect_caps(oracle, boxes, host, reqs)
    biggest = max(w.shape[0] for w in want)
    ix = make_index(boxes)
    res, st = S.run(ix, reqs, mode=mode, arena_pairs=max(4096, biggest // 3 + 1), chunk_queries=64)
    assert st["relaunches"] > 0
    for i, (r, w) in enumerate(zip(res, want)):
        assert np.array_equal(canon(r["pairs"]), w), i


@pytest.mark.parametrize("mode,policy", [("batched", "sah"), ("plain", "sah"),
                                         ("batched", "never"), ("plain", "never")])
def test_updates_interleaved_each_request_sees_its_epoch(oracle, scene, mode, policy):
    # Under the SAH policy these 5%-of-the-boxes update batches rebuild the
    # tree; under RTS_REBUILD_NEVER every update is a refit, so both the
    # rebuild and the refit path of an update run between batches.
    rng, means, boxes0 = scene
    log(f"serve with updates mode={mode} policy={policy}")
    boxes0 = boxes0[:100_000].copy()
    r2 = np.random.default_rng(4)
    reqs, host = make_trace(r2, 60, means, 2000, rate_hz=3000.0)
    # 4 update batches spread over the trace
    span = reqs[-1].arrival_s
    ups, states = [], [boxes0.copy()]
    cur = boxes0.copy()
    for k in range(4):
        ids = r2.choice(cur.shape[0], 5000, replace=False).astype(np.int64)
        shift = r2.uniform(-0.01, 0.01, size=(5000, 2)).astype(np.float32)
        nb = cur[ids].copy()
        nb[:, [0, 2]] += shift[:, :1]
        nb[:, [1, 3]] += shift[:, 1:]
        cur = cur.copy()
        cur[ids] = nb
        states.append(cur)
        ups.append(S.UpdateBatch(ids=dev(ids), boxes=dev(nb), arrival_s=span * (k + 1) / 5))
    trees = [oracle.RTree(s) for s in states]
    for (q, pred), r in zip(host, reqs):
        r.collect = max(t.count(int(pred), q) for t in trees) + 8
    never = policy == "never"
    ix = make_index(boxes0, **({"rebuild_policy": N.RTS_REBUILD_NEVER} if never else {}))
    ix.Build()
    ix.Sync()
    builds0 = ix.stats()["builds"]
    res, st = S.run(ix, reqs, ups, mode=mode, slots=3)
    assert st["updates"] == 4
    stats = ix.stats()
    if never:
        assert stats["refits"] >= 4 and stats["builds"] == builds0, stats
    else:
        assert stats["builds"] >= builds0 + 4, stats
    epochs = set()
    for i, ((q, pred), r) in enumerate(zip(host, res)):
        e = r["epoch"]
        epochs.add(e)
        assert np.array_equal(canon(r["pairs"]), trees[e].pairs(int(pred), q)), (i, e)
    assert len(epochs) >= 2
    # The caller's index ends with every update applied.
    q, pred = host[0]
    got, _ = collect(ix, pred, dev(q))
    assert np.array_equal(got, trees[-1].pairs(int(pred), q))


@pytest.mark.parametrize("sort_min_tiles", [1, 4096])
def test_giant_queries_beside_small_ones(oracle, sort_min_tiles):
    # Queries of 1.2M, ~300k and ~70k results (the write pass takes them
    # first, largest first) in batches with thousands of small queries,
    # with and without the Morton order of the wave tiles.
    rng = np.random.default_rng(5)
    n = 1_200_000
    lo = rng.random((n, 2), dtype=np.float32)
    boxes = np.concatenate([lo, lo + np.float32(1e-4)], axis=1)
    log(f"serve giants sort_min_tiles={sort_min_tiles}")

    def env(k, side):
        c = rng.random((k, 2), dtype=np.float32) * np.float32(1 - side)
        return np.concatenate([c, c + np.float32(side)], axis=1).astype(np.float32)

    host = [
        (np.array([[-1, -1, 2, 2], [-1, -1, 2, 2]], np.float32), P.kIntersects),
        (env(3000, 2e-3), P.kIntersects),
        (env(3, 0.5), P.kWithin),
        (env(2000, 3e-3), P.kWithin),
        (env(4, 0.24), P.kIntersects),
        (rng.random((5000, 2), dtype=np.float32), P.kContains),
        (np.array([[-1, -1, 2, 2]], np.float32), P.kWithin),
    ]
    reqs = [S.Request(queries=dev(q), predicate=p, arrival_s=0.0, collect=0) for q, p in host]
    want = _collect_caps(oracle, boxes, host, reqs)
    assert max(w.shape[0] for w in want) >= 1 << 20
    ix = make_index(boxes)
    res, st = S.run(ix, reqs, mode="batched", slots=3, sort_min_tiles=sort_min_tiles)
    assert st["requests"] == len(host)
    for i, (r, w) in enumerate(zip(res, want)):
        assert r["count"] == w.shape[0], i
        assert not r["truncated"]
        assert np.array_equal(canon(r["pairs"]), w), i


def test_empty_and_single_query_requests(oracle, scene):
    rng, means, boxes = scene
    log("serve tiny requests")
    q1 = np.float32([[0.5, 0.5, 0.51, 0.51]])
    reqs = [S.Request(queries=dev(np.zeros((0, 4), np.float32)), collect=1),
            S.Request(queries=dev(q1), predicate=P.kIntersects, collect=100000),
            S.Request(queries=dev(np.float32([[0.5, 0.5]])), predicate=P.kContains, collect=1000)]
    ix = make_index(boxes)
    res, _ = S.run(ix, reqs)
    assert res[0]["count"] == 0
    assert np.array_equal(canon(res[1]["pairs"]), oracle.bf_pairs(oracle.INTERSECTS, boxes, q1))
    assert np.array_equal(canon(res[2]["pairs"]),
                          oracle.bf_pairs(oracle.CONTAINS, boxes, np.float32([[0.5, 0.5]])))


@pytest.mark.parametrize("mode", ["batched", "plain"])
def test_short_output_keeps_the_first_pairs_in_query_order(oracle, scene, mode):
    # `out` holds half of a one-chunk request's results: the count is still
    # the whole count, `truncated` is set, and `out` holds the pairs of the
    # request's first queries in query order, the last of them possibly in
    # part. (The batched path gathers each query's pairs from the start of
    # its slots; the plain path copies one contiguous range.)
    rng, means, boxes = scene
    log(f"serve truncated output mode={mode}")
    q = W.mixture_boxes(np.random.default_rng(11), means, 3000, np.log(5e-3))
    want = oracle.RTree(boxes).pairs(oracle.INTERSECTS, q)  # (query, geom), sorted
    cap = want.shape[0] // 2
    assert cap > 0
    ix = make_index(boxes)
    res, _ = S.run(ix, [S.Request(queries=dev(q), predicate=P.kIntersects, collect=cap)],
                   mode=mode)
    r = res[0]
    assert r["count"] == want.shape[0] and r["truncated"]
    assert r["pairs"].shape[0] == cap
    qids = r["pairs"].cpu().numpy()[:, 1].astype(np.int64)
    assert np.all(np.diff(qids) >= 0)
    got = canon(r["pairs"])
    last = got[-1, 0]
    assert np.array_equal(got[got[:, 0] < last], want[want[:, 0] < last])
    part = {tuple(x) for x in got[got[:, 0] == last]}
    assert part <= {tuple(x) for x in want[want[:, 0] == last]}


@pytest.mark.parametrize("mode", ["batched", "plain"])
def test_empty_index_serves_empty_results(mode):
    # No box was ever inserted: every request completes with no results and
    # no pairs, in both modes (the batched count pass writes zero counts).
    log(f"serve on an empty index mode={mode}")
    ix = R.SpatialIndex()
    ix.Init(R.Config())
    q4 = dev(np.float32([[0.1, 0.1, 0.9, 0.9]] * 300))
    q2 = dev(np.float32([[0.5, 0.5]] * 70))
    reqs = [S.Request(queries=q4, predicate=P.kIntersects, collect=64),
            S.Request(queries=q2, predicate=P.kContains, collect=64),
            S.Request(queries=q4, predicate=P.kWithin, collect=64)]
    res, st = S.run(ix, reqs, mode=mode, hash_results=1)
    assert st["requests"] == 3
    for r in res:
        assert r["count"] == 0 and r["hash"] == 0 and not r["truncated"]
        assert r["pairs"].shape[0] == 0


def test_graph_replay_is_refused(scene):
    # hipGraph replay of a batch did not complete on the MI355X and its cause
    # is not established: the option is refused before anything runs.
    rng, means, boxes = scene
    ix = make_index(boxes[:1000])
    reqs = [S.Request(queries=dev(np.float32([[0.5, 0.5, 0.6, 0.6]])), collect=8)]
    with pytest.raises(RuntimeError, match="use_graphs"):
        S.run(ix, reqs, mode="batched", use_graphs=1)


@pytest.mark.parametrize("mode", ["batched", "plain"])
def test_segment_requests_beside_envelopes_and_points(oracle, scene, mode):
    # Segment requests (one count and one write launch each inside their
    # batch) mixed with tiled envelope and point requests; a segment request
    # under RTS_CONTAINS matches nothing, as the plain call.
    rng, means, boxes = scene
    log(f"serve segments mode={mode}")
    r = np.random.default_rng(11)
    host = []
    for i in range(24):
        n = int(r.integers(1, 400))
        kind = i % 4
        if kind == 0:
            q, pred, seg = W.mixture_boxes(r, means, n, np.log(4e-3)), P.kIntersects, True
        elif kind == 1:
            q, pred, seg = W.mixture_boxes(r, means, n, np.log(4e-3)), P.kWithin, False
        elif kind == 2:
            q, pred, seg = W.mixture_boxes(r, means, n, np.log(4e-3))[:, :2].copy(), P.kContains, False
        else:
            q, pred, seg = W.mixture_boxes(r, means, n, np.log(4e-3)), P.kIntersects, False
        host.append((q, pred, seg))
    host.append((W.mixture_boxes(r, means, 50, np.log(4e-3)), P.kContains, True))
    want = [oracle.bf_pairs(int(p), boxes, q, lines=s) for q, p, s in host]
    reqs = [S.Request(queries=dev(q), predicate=p, arrival_s=0.001 * i, collect=w.shape[0] + 8,
                      segments=s) for i, ((q, p, s), w) in enumerate(zip(host, want))]
    ix = make_index(boxes)
    res, st = S.run(ix, reqs, mode=mode, slots=3, hash_results=1)
    assert st["requests"] == len(host)
    for i, (rr, w) in enumerate(zip(res, want)):
        assert rr["count"] == w.shape[0], i
        assert np.array_equal(canon(rr["pairs"]), w), i
    assert want[-1].shape[0] == 0 and any(w.shape[0] for w in want[:-1])


@pytest.mark.parametrize("mode", ["batched", "plain"])
def test_double_index_requests(oracle, mode):
    # An exact double index: every request (envelopes under each predicate,
    # points, segments) is walked in double, with coordinates within 1e-13 of
    # data coordinates where the float bounds cannot decide.
    log(f"serve double index mode={mode}")
    rng = np.random.default_rng(641)
    lo = rng.random((40_000, 2))
    boxes = np.concatenate([lo, lo + rng.random((40_000, 2)) * 0.01], axis=1)

    def near(n):
        src = boxes[rng.integers(0, boxes.shape[0], n)]
        q = src + rng.choice([-1e-13, 0.0, 1e-13], size=(n, 4))
        q[:, 2:] = np.maximum(q[:, 2:], q[:, :2]) + rng.random((n, 2)) * 0.02
        return q

    host = [(near(300), P.kIntersects, False), (near(200), P.kWithin, False),
            (near(100), P.kContains, False),
            (boxes[rng.integers(0, boxes.shape[0], 300), :2] + rng.choice([-1e-13, 0.0, 1e-13], size=(300, 2)),
             P.kContains, False),
            (near(150), P.kIntersects, True)]
    want = [oracle.bf_pairs64(int(p), boxes, q, lines=s) for q, p, s in host]
    reqs = [S.Request(queries=dev(np.ascontiguousarray(q)), predicate=p, arrival_s=0.0,
                      collect=w.shape[0] + 8, segments=s) for (q, p, s), w in zip(host, want)]
    ix = make_index(boxes)
    res, st = S.run(ix, reqs, mode=mode, slots=2)
    for i, (rr, w) in enumerate(zip(res, want)):
        assert rr["count"] == w.shape[0], i
        assert np.array_equal(canon(rr["pairs"]), w), i
    assert all(w.shape[0] for w in want)


def test_double_as_float_index_is_refused():
    ix = make_index(np.array([[0.0, 0.0, 1.0, 1.0]]), coord=N.RTS_COORD_F64_AS_F32)
    with pytest.raises(N.RtsError, match="RTS_COORD_F64_AS_F32"):
        S.run(ix, [S.Request(queries=dev(np.array([[0.5, 0.5]])), predicate=P.kContains)])


def test_double_index_updates_each_request_sees_its_epoch(oracle):
    # Update batches on a double index (rts_update_d records) between
    # requests: each request equals the exact double join on its epoch.
    log("serve double index with updates")
    rng = np.random.default_rng(642)
    lo = rng.random((20_000, 2))
    b0 = np.concatenate([lo, lo + rng.random((20_000, 2)) * 0.01], axis=1)
    states, ups = [b0], []
    cur = b0
    for k in range(2):
        ids = rng.choice(cur.shape[0], 2000, replace=False).astype(np.int64)
        nb = cur[ids] + rng.uniform(-0.005, 0.005, size=(2000, 1)) * np.array([1, 0, 1, 0]) \
            + rng.uniform(-0.005, 0.005, size=(2000, 1)) * np.array([0, 1, 0, 1])
        cur = cur.copy()
        cur[ids] = nb
        states.append(cur)
        ups.append(S.UpdateBatch(ids=dev(ids), boxes=dev(nb), arrival_s=0.02 * (k + 1)))
    host = []
    for i in range(12):
        q = rng.random((200, 4))
        q[:, 2:] = q[:, :2] + rng.random((200, 2)) * 0.02
        host.append((q, (P.kIntersects, P.kWithin)[i % 2]))
    reqs = [S.Request(queries=dev(q), predicate=p, arrival_s=0.006 * i,
                      collect=max(oracle.bf_count64(int(p), s, q) for s in states) + 8)
            for i, (q, p) in enumerate(host)]
    ix = make_index(b0)
    res, st = S.run(ix, reqs, ups, mode="batched", slots=2)
    assert st["updates"] == 2
    epochs = set()
    for i, ((q, p), rr) in enumerate(zip(host, res)):
        e = rr["epoch"]
        epochs.add(e)
        assert np.array_equal(canon(rr["pairs"]), oracle.bf_pairs64(int(p), states[e], q)), (i, e)
    assert len(epochs) >= 2
