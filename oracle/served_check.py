"""Checks a served run of the mixed stream (rtspatial_amd.workloads.ServedStream)
against the oracle: every request's result count and pair checksum must equal
the oracle's on the index state of the request's epoch (the boxes after the
first `epoch` update batches).

TEST INFRASTRUCTURE ONLY: called by tests/ and by bench.py's verification
stage (outside any timed region), never by the rtspatial_amd package.

The oracle is an STR R-tree over the stream's boxes, refitted in place as the
update batches are applied in order; each (epoch, predicate) evaluates the
union of its requests' pool slices once, and a request's checksum is
request_checksum() over its slice (query ids local to the request, as the
served path reports them)."""
import time

import numpy as np

from . import oracle as O


def check(stream, results, threads=0, log=None):
    """Returns a dict: requests and epochs checked, queries evaluated,
    mismatches (up to 16, with both answers) and oracle seconds."""
    t0 = time.perf_counter()
    tree = O.RTree(stream.boxes, threads=threads)
    by_epoch = {}
    for i, r in enumerate(results):
        by_epoch.setdefault(int(r["epoch"]), []).append(i)
    applied = 0
    mism = []
    evaluated = 0
    checked = 0
    for e in sorted(by_epoch):
        while applied < e:
            ids, nb = stream.updates[applied]
            tree.update(ids.astype(np.uint64), nb)
            applied += 1
        for kind in (0, 1, 2):
            idx = [i for i in by_epoch[e] if int(stream.kinds[i]) == kind]
            if not idx:
                continue
            pool = stream.pt_pool if kind == 0 else stream.env_pool
            cover = np.zeros(stream.pool_n + 1, np.int64)
            for i in idx:
                cover[int(stream.offsets[i])] += 1
                cover[int(stream.offsets[i]) + int(stream.sizes[i])] -= 1
            qidx = np.nonzero(np.cumsum(cover[:-1]) > 0)[0]
            _, cnt, gs = tree.checksums(stream.predicate(idx[0]), pool[qidx], threads)
            evaluated += qidx.shape[0]
            full_cnt = np.zeros(stream.pool_n, np.uint64)
            full_gs = np.zeros(stream.pool_n, np.uint64)
            full_cnt[qidx] = cnt
            full_gs[qidx] = gs
            for i in idx:
                o, n = int(stream.offsets[i]), int(stream.sizes[i])
                want = (int(full_cnt[o:o + n].sum(dtype=np.uint64)),
                        O.request_checksum(full_gs[o:o + n]))
                got = (int(results[i]["count"]), int(results[i]["hash"]))
                checked += 1
                if got != want or results[i].get("truncated"):
                    if len(mism) < 16:
                        mism.append({"request": i, "epoch": e, "kind": kind, "queries": n,
                                     "got": list(got), "oracle": list(want)})
                    else:
                        mism.append(None)
        if log:
            log(f"oracle check: epoch {e} done ({checked} requests so far)")
    return {"requests": checked, "epochs": len(by_epoch), "updates_applied": applied,
            "queries_evaluated": int(evaluated), "oracle_equal": not mism,
            "mismatches": [m for m in mism if m is not None], "n_mismatches": len(mism),
            "oracle_s": round(time.perf_counter() - t0, 2)}
