"""Batch-structured insert (reference spatial_index.cuh:142-196: one GAS per
Insert call under an IAS, merged by Optimize, :766-888). Here every Insert
appends to one set of boxes and the tree is rebuilt once before the next
query, so results must not depend on how the boxes were batched: after every
batch, queries over the boxes inserted so far equal the oracle's."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():
    pytest.skip("no GPU", allow_module_level=True)

import rtspatial_amd as R  # noqa: E402
from gpu_util import collect, count, dev, log  # noqa: E402

P = R.Predicate


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_insert_batches_query_between(oracle, dtype):
    from rtspatial_amd import workloads as W
    log(f"insert batches, {np.dtype(dtype).name}")
    rng = np.random.default_rng(70)
    means = W.mixture_means(rng)
    sizes = [1, 7, 64, 1000, 3, 50_000, 12_345, 2, 30_000, 777] * 2
    total = sum(sizes)
    boxes = W.mixture_boxes(rng, means, total, np.log(2e-3), dtype=dtype)
    q = W.mixture_boxes(rng, means, 800, np.log(1e-2), dtype=dtype)
    pts = q[:, :2].copy()
    ix = R.SpatialIndex(torch.float64 if dtype == np.float64 else torch.float32)
    ix.Init(R.Config(max_geometries=1000))  # grows on demand
    alive = np.ones(total, np.uint8)
    done = 0
    f32 = dtype == np.float32
    for k, n in enumerate(sizes):
        ix.Insert(dev(boxes[done:done + n]))
        done += n
        assert ix.size() == done
        if k % 4 == 3:
            ix.Optimize()
        if k % 5 == 4:  # delete some ids of earlier batches
            dead = rng.choice(done, min(500, done // 3), replace=False).astype(np.int64)
            ix.Delete(dev(dead))
            alive[dead] = 0
        b, a = boxes[:done], alive[:done]
        want_i = (oracle.bf_pairs if f32 else oracle.bf_pairs64)(oracle.INTERSECTS, b, q, alive=a)
        got, n_i = collect(ix, P.kIntersects, q)
        assert n_i == want_i.shape[0] and np.array_equal(got, want_i), f"batch {k}"
        want_p = (oracle.bf_count if f32 else oracle.bf_count64)(oracle.CONTAINS, b, pts, alive=a)
        assert count(ix, P.kContains, pts) == want_p, f"batch {k}"
    st = ix.stats()
    assert st["n_boxes"] == total and st["n_tree"] == total and st["builds"] >= 1
    m = ix.memory()
    assert m["geometries"] >= total * (16 if f32 else 48) and m["aabbs"] >= total * 16
    assert m["total"] == m["geometries"] + m["aabbs"] + m["bvh"] + 256
    ix.PrintMemoryUsage()
    ix.Sync()


def test_optimize_and_memory_on_empty_index():
    log("optimize on empty index")
    ix = R.SpatialIndex()
    ix.Init(R.Config())
    ix.Optimize()
    m = ix.memory()
    assert m["bvh"] == 0 and m["aabbs"] == 0
    ix.Sync()
