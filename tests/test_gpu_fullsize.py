"""Parity at the benchmark configurations' full sizes.

  c2  1M boxes x 1M query envelopes, within (data box inside the query):
      the full sorted (query_id, box_id) pair set equals the oracle's.
  c3  10M skewed boxes x 1M query envelopes, intersects (~1e10 pairs, all
      collected on the device): every query's count and ghash checksum
      (rts_pairs_checksum over the collected pairs) equal the oracle's, and
      the full pair sets of 1000 sampled queries are equal.
  c4  10M uniform boxes, 20 rounds of 1M deletes, 1M inserts, 1M updates and
      1M point queries, on three indexes side by side (SAH policy; refit-only
      with a large delta tree; rebuild always): after every round each
      index's full sorted pair set equals the oracle's.
  c5  the served mixed stream over the 10M-box index with 100k-box update
      batches at 10 Hz: every request's count and pair checksum equal the
      oracle's on the index state of the request's epoch.

The oracle is the STR R-tree of oracle/ (pinned against brute force and the
reference's known answers by tests/test_oracle.py)."""
import math
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():
    pytest.skip("no GPU", allow_module_level=True)

import rtspatial_amd as R  # noqa: E402
from rtspatial_amd import _native as N  # noqa: E402
from rtspatial_amd import serve as S  # noqa: E402
from rtspatial_amd import workloads as W  # noqa: E402
from gpu_util import canon, collect, count, dev, log  # noqa: E402

P = R.Predicate


def _index(boxes, **cfg):
    ix = R.SpatialIndex()
    ix.Init(R.Config(max_geometries=boxes.shape[0], **cfg))
    ix.Insert(boxes if isinstance(boxes, torch.Tensor) else dev(boxes))
    ix.Build()
    ix.Sync()
    return ix


def gpu_checksums(queue, n_pairs, n_queries):
    """Per-query counts and ghash sums of the first n_pairs pairs of a queue,
    computed on the device by rts_pairs_checksum."""
    cnt = torch.empty(n_queries, dtype=torch.int64, device="cuda")
    gs = torch.empty(n_queries, dtype=torch.int64, device="cuda")
    N.check(N.lib().rts_pairs_checksum(queue.pairs().data_ptr(), n_pairs, n_queries,
                                       cnt.data_ptr(), gs.data_ptr(), None), "pairs_checksum")
    torch.cuda.synchronize()
    return cnt.cpu().numpy().view(np.uint64), gs.cpu().numpy().view(np.uint64)


def test_c2_within_1M_x_1M_full_pair_set(oracle):
    t0 = time.time()
    boxes, queries = W.config2()
    log(f"c2: 1M boxes x 1M within queries (generated in {time.time() - t0:.1f} s)")
    ix = _index(boxes)
    got, n = collect(ix, P.kWithin, queries)
    log(f"c2: {n} pairs from the GPU")
    want = oracle.RTree(boxes).pairs(oracle.WITHIN, queries)
    assert n == want.shape[0] and n > 10_000_000
    assert np.array_equal(got, want)
    ix.Sync()
    log(f"c2: equal ({time.time() - t0:.1f} s)")


def test_c3_intersects_10M_x_1M_checksums_and_sample(oracle):
    t0 = time.time()
    boxes, queries = W.config3()
    nq = queries.shape[0]
    log(f"c3: 10M boxes x 1M intersects queries (generated in {time.time() - t0:.1f} s)")
    ix = _index(boxes)
    q = dev(queries)
    total = count(ix, P.kIntersects, q)
    log(f"c3: {total} pairs counted; collecting them all")
    queue = R.Queue(total + 64)
    ix.Query(P.kIntersects, q, queue)
    n = queue.size()
    assert n == total
    cnt, gs = gpu_checksums(queue, n, nq)
    log(f"c3: GPU checksums done ({time.time() - t0:.1f} s); oracle R-tree")
    tree = oracle.RTree(boxes)
    wtot, wcnt, wgs = tree.checksums(oracle.INTERSECTS, queries)
    log(f"c3: oracle done ({time.time() - t0:.1f} s)")
    assert wtot == total
    bad = np.nonzero((cnt != wcnt) | (gs != wgs))[0]
    assert bad.size == 0, f"{bad.size} queries differ, first {bad[:8].tolist()}"
    # Full pair sets of every 1000th query, taken from the collected pairs.
    pairs = queue.pairs()[:n]
    sel = []
    step = 1 << 30
    for s in range(0, n, step):
        chunk = pairs[s:s + step]
        keep = (chunk[:, 1] % 1000) == 0
        sel.append(chunk[keep].clone())
    sample = torch.cat(sel)
    del queue, pairs, sel
    got = canon(sample)
    got[:, 0] //= 1000
    want = tree.pairs(oracle.INTERSECTS, queries[::1000])
    assert got.shape[0] > 1_000_000
    assert np.array_equal(got, want)
    ix.Sync()
    log(f"c3: equal ({time.time() - t0:.1f} s)")


def test_c4_mutable_10M_20_rounds_three_policies(oracle):
    t0 = time.time()
    m = W.MutationRounds()
    log(f"c4: 10M boxes, 20 rounds (generated in {time.time() - t0:.1f} s)")
    pols = {"sah": dict(),
            "never": dict(rebuild_policy=N.RTS_REBUILD_NEVER, insert_delta_pct=100),
            "always": dict(rebuild_policy=N.RTS_REBUILD_ALWAYS)}
    init = dev(m.initial)
    idx = {k: _index(init, **c) for k, c in pols.items()}
    del init
    base = {k: ix.stats() for k, ix in idx.items()}
    for rnd in range(20):
        del_ids, ins, upd_ids, upd, pts = m.next_round()
        d_del, d_ins, d_upd, d_pts = dev(del_ids), dev(ins), (dev(upd_ids), dev(upd)), dev(pts)
        got = {}
        for k, ix in idx.items():
            ix.Delete(d_del)
            ix.Insert(d_ins)
            ix.Update(d_upd)
            got[k], _ = collect(ix, P.kContains, d_pts)
        want = oracle.RTree(m.boxes, alive=m.alive.astype(np.uint8)).pairs(oracle.CONTAINS, pts)
        for k in idx:
            assert np.array_equal(got[k], want), (rnd, k)
        log(f"c4: round {rnd}: {want.shape[0]} pairs equal on all three indexes "
            f"({time.time() - t0:.1f} s)")
    st = {k: ix.stats() for k, ix in idx.items()}
    # Refit-only with a delta tree refits and builds fewer full trees than
    # the rebuild-always index, which rebuilds every round.
    assert st["never"]["refits"] > base["never"]["refits"]
    assert st["never"]["builds"] < st["always"]["builds"]
    assert st["always"]["builds"] >= base["always"]["builds"] + 20
    for ix in idx.values():
        ix.Sync()


def test_c5_served_stream_10M_checksums_per_epoch(oracle):
    from oracle import served_check as SC

    t0 = time.time()
    s = W.ServedStream(n_requests=150, n_update_batches=40)
    log(f"c5: 10M-box index, 150 requests, {int(s.sizes.sum())} queries "
        f"(generated in {time.time() - t0:.1f} s)")
    ix = _index(s.boxes)
    env, pts = dev(s.env_pool), dev(s.pt_pool)
    preds = [P.kContains, P.kWithin, P.kIntersects]
    rate = 150.0  # requests per second: the run spans ~1 s, so ~10 update batches arrive
    t = s.arrivals(rate)
    reqs = []
    for i in range(len(t)):
        o, n = int(s.offsets[i]), int(s.sizes[i])
        q = pts[o:o + n] if s.kinds[i] == 0 else env[o:o + n]
        reqs.append(S.Request(queries=q, predicate=preds[int(s.kinds[i])], arrival_s=float(t[i])))
    ups = [S.UpdateBatch(ids=dev(s.updates[k][0]), boxes=dev(s.updates[k][1]), arrival_s=a)
           for k, a in enumerate(s.update_times(float(t[-1]) + 0.5))]
    res, st = S.run(ix, reqs, ups, mode="batched", hash_results=1, slots=3,
                    arena_pairs=1_500_000_000, chunk_queries=128 << 10, head_share_pct=75)
    log(f"c5: served {st['queries']} queries, {st['results']} pairs, {st['updates']} updates, "
        f"{st['batches']} batches, {st['relaunches']} relaunches ({time.time() - t0:.1f} s)")
    assert st["updates"] >= 3
    rep = SC.check(s, res, log=log)
    log(f"c5: oracle check {rep['requests']} requests over {rep['epochs']} epochs "
        f"({rep['oracle_s']} s)")
    assert rep["oracle_equal"], rep["mismatches"]
    assert rep["epochs"] >= 3 and rep["requests"] == 150
    assert any(int(s.sizes[i]) > 200_000 for i in range(150))
    ix.Sync()
