"""The served-run checker (oracle/served_check.py) against brute force: for
a small mixed stream, per-request answers computed by the brute-force join on
each request's epoch state must pass the check, and any altered answer, a
wrong epoch or an update applied out of order must fail it."""
import numpy as np

from rtspatial_amd import workloads as W


def _brute_results(oracle, s, epochs):
    res = []
    cur = s.boxes.copy()
    states = [cur.copy()]
    for ids, nb in s.updates:
        cur = cur.copy()
        cur[ids] = nb
        states.append(cur)
    for i, e in enumerate(epochs):
        q = s.queries(i)
        pairs = oracle.bf_pairs(s.predicate(i), states[e], q)
        cnt, gs = oracle.pair_checksums(pairs, q.shape[0])
        res.append({"count": int(cnt.sum()), "hash": oracle.request_checksum(gs),
                    "epoch": e, "truncated": 0})
    return res


def test_served_check_against_brute_force(oracle):
    from oracle import served_check as SC

    s = W.ServedStream(n_boxes=20000, n_requests=40, max_request=300, pool_n=4096,
                       n_update_batches=4, update_size=3000)
    rng = np.random.default_rng(1)
    epochs = np.sort(rng.integers(0, 5, size=40))
    res = _brute_results(oracle, s, epochs)
    assert sum(r["count"] for r in res) > 0
    rep = SC.check(s, res)
    assert rep["oracle_equal"] and rep["requests"] == 40 and rep["updates_applied"] == int(epochs.max())
    # An altered count, an altered hash and a wrong epoch are each caught.
    for field, delta in (("count", 1), ("hash", 1)):
        bad = [dict(r) for r in res]
        k = next(i for i, r in enumerate(res) if r["count"] > 0)
        bad[k][field] += delta
        rep = SC.check(s, bad)
        assert not rep["oracle_equal"] and rep["mismatches"][0]["request"] == k
    bad = [dict(r) for r in res]
    k = next(i for i, r in enumerate(res) if r["epoch"] > 0 and r["count"] > 0
             and s.kinds[i] != 0)
    bad[k]["epoch"] -= 1
    assert not SC.check(s, bad)["oracle_equal"]
