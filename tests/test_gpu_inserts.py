"""Batch-structured insert (reference spatial_index.cuh:142-196: one GAS per
Insert call under an IAS, merged by Optimize, :766-888). Here every Insert
appends to one set of boxes and the tree is rebuilt once before the next
query, so results must not depend on how the boxes were batched: after every
batch, queries over the boxes inserted so far equal the oracle's."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():
    pytest.skip("no GPU", allow_module_level=True)

import rtspatial_amd as R  # noqa: E402
from gpu_util import collect, count, dev, log  # noqa: E402

P = R.Predicate


@pytest.mark.parametrize("dtype,delta_pct", [(np.float32, 2), (np.float32, 0), (np.float32, 50),
                                              (np.float64, 50)])
def test_insert_batches_query_between(oracle, dtype, delta_pct):
    # delta_pct: inserts since the last full build go to a small second tree
    # while they are at most this share of the main tree (0: always rebuild).
    from rtspatial_amd import workloads as W
    log(f"insert batches, {np.dtype(dtype).name}, delta {delta_pct}%")
    rng = np.random.default_rng(70)
    means = W.mixture_means(rng)
    sizes = [1, 7, 64, 1000, 3, 50_000, 12_345, 2, 30_000, 777] * 2
    total = sum(sizes)
    boxes = W.mixture_boxes(rng, means, total, np.log(2e-3), dtype=dtype)
    q = W.mixture_boxes(rng, means, 800, np.log(1e-2), dtype=dtype)
    pts = q[:, :2].copy()
    ix = R.SpatialIndex(torch.float64 if dtype == np.float64 else torch.float32)
    ix.Init(R.Config(max_geometries=1000, insert_delta_pct=delta_pct))  # grows on demand
    cur = boxes.copy()  # the boxes as the index holds them
    alive = np.ones(total, np.uint8)
    done = 0
    f32 = dtype == np.float32
    deltas = []
    for k, n in enumerate(sizes):
        ix.Insert(dev(boxes[done:done + n]))
        done += n
        assert ix.size() == done
        if k % 4 == 3:
            ix.Optimize()
        if k % 5 == 4:  # delete and update ids, the newest ones included
            dead = rng.choice(done, min(500, done // 3), replace=False).astype(np.int64)
            ix.Delete(dev(dead))
            alive[dead] = 0
        if k % 3 == 2 and done > 100:
            ids = np.unique(np.concatenate([np.arange(done - 20, done),
                                            rng.choice(done, 200, replace=False)])).astype(np.int64)
            new = cur[rng.integers(0, done, ids.shape[0])] + dtype(1e-3)
            ix.Update((dev(ids), dev(new)))
            cur[ids] = new
            alive[ids] = 1
        b, a = cur[:done], alive[:done]
        want_i = (oracle.bf_pairs if f32 else oracle.bf_pairs64)(oracle.INTERSECTS, b, q, alive=a)
        got, n_i = collect(ix, P.kIntersects, q)
        assert n_i == want_i.shape[0] and np.array_equal(got, want_i), f"batch {k}"
        want_p = (oracle.bf_count if f32 else oracle.bf_count64)(oracle.CONTAINS, b, pts, alive=a)
        assert count(ix, P.kContains, pts) == want_p, f"batch {k}"
        deltas.append(ix.stats()["delta_boxes"])
    if delta_pct == 0:
        assert max(deltas) == 0
    else:
        assert max(deltas) > 0  # some batches went to the delta tree
    st = ix.stats()
    assert st["n_boxes"] == total and st["n_tree"] == total and st["builds"] >= 1
    m = ix.memory()
    assert m["geometries"] >= total * (16 if f32 else 48) and m["aabbs"] >= total * 16
    assert m["total"] == m["geometries"] + m["aabbs"] + m["bvh"] + 256
    ix.PrintMemoryUsage()
    ix.Sync()


def test_delta_tree_merged_for_served_path_and_tree_view(oracle):
    # The served path and caller-compiled handlers walk one tree: a delta tree
    # is merged into the main one first, and results stay exact.
    from rtspatial_amd import serve as S
    from rtspatial_amd import workloads as W
    log("delta tree merged for one-tree callers")
    rng = np.random.default_rng(71)
    means = W.mixture_means(rng)
    b = W.mixture_boxes(rng, means, 100_000, np.log(2e-3))
    extra = W.mixture_boxes(rng, means, 1_000, np.log(2e-3))
    q = W.mixture_boxes(rng, means, 500, np.log(1e-2))
    ix = R.SpatialIndex()
    ix.Init(R.Config(insert_delta_pct=5))
    ix.Insert(dev(b))
    ix.Build()
    ix.Insert(dev(extra))
    ix.Build()
    assert ix.stats()["delta_boxes"] == 1_000
    allb = np.concatenate([b, extra])
    want = oracle.bf_count(oracle.INTERSECTS, allb, q)
    assert count(ix, P.kIntersects, q) == want  # both trees walked
    res, st = S.run(ix, [S.Request(queries=dev(q), predicate=P.kIntersects)], mode="batched")
    assert ix.stats()["delta_boxes"] == 0 and int(res[0]["count"]) == want
    ix.Insert(dev(extra[:10]))
    ix.Build()
    assert ix.stats()["delta_boxes"] == 10
    ix.Optimize()  # merges, as the reference's Optimize merges its per-Insert BVHs
    assert ix.stats()["delta_boxes"] == 0
    ix.Insert(dev(extra[10:20]))
    ix.Build()
    assert ix.stats()["delta_boxes"] == 10
    v = R._native.rts_tree_view()
    N = R._native
    N.check(N.lib().rts_index_tree_view(ix._handle(), None, ctypes.byref(v)), "tree view")
    assert v.n == 101_020 and ix.stats()["delta_boxes"] == 0
    ix.Sync()


def test_optimize_and_memory_on_empty_index():
    log("optimize on empty index")
    ix = R.SpatialIndex()
    ix.Init(R.Config())
    ix.Optimize()
    m = ix.memory()
    assert m["bvh"] == 0 and m["aabbs"] == 0
    ix.Sync()
