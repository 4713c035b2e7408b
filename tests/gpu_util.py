"""Helpers shared by the GPU tests: run the HIP path through the package's
C-ABI binding and canonicalise its output for comparison with the oracle."""
import sys

import numpy as np
import torch

import rtspatial_amd as R


def log(msg):
    print(f"[gpu-test] {msg}", flush=True)
    sys.stdout.flush()


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def pinned(a):
    t = torch.from_numpy(np.ascontiguousarray(a)).pin_memory()
    return t


def make_index(boxes, batches=1, **cfg):
    """An index over `boxes` (float32 or, for a double index, float64)."""
    b = dev(boxes) if isinstance(boxes, np.ndarray) else boxes
    ix = R.SpatialIndex(torch.float64 if b.dtype == torch.float64 else torch.float32)
    ix.Init(R.Config(**cfg))
    n = b.shape[0]
    step = (n + batches - 1) // batches if batches > 1 else n
    for s in range(0, n, max(step, 1)):
        ix.Insert(b[s:s + step])
    return ix


def count(ix, pred, queries):
    q = dev(queries) if isinstance(queries, np.ndarray) else queries
    c = R.SharedValue()
    c.set(0)
    ix.Query(pred, q, c)
    return c.get()


def canon(pairs_geom_query):
    """(geom_id, query_id) int32 tensor -> (query_id, geom_id) uint32 numpy
    array sorted by (query, geom)."""
    p = pairs_geom_query.cpu().numpy().astype(np.uint32)
    qg = np.ascontiguousarray(p[:, ::-1])
    key = qg[:, 0].astype(np.uint64) << np.uint64(32) | qg[:, 1].astype(np.uint64)
    order = np.argsort(key, kind="stable")
    return qg[order]


def collect(ix, pred, queries, capacity=None):
    """Runs a collecting query; sizes the queue with a counting run first
    unless `capacity` is given. Returns (sorted (query, geom) pairs, counter)."""
    q = dev(queries) if isinstance(queries, np.ndarray) else queries
    if capacity is None:
        capacity = count(ix, pred, q) + 1
    queue = R.Queue(capacity)
    ix.Query(pred, q, queue)
    n = queue.size()
    return canon(queue.pairs()), n
