"""The built tree's depth on the deepest inputs 30-bit Morton keys allow (a
chain of one-bit keys above a large run of equal keys), read back from the
device through rts_index_tree_view, stays within the bound the per-lane
stack is sized for, and per-lane point queries over it are exact."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():
    pytest.skip("no GPU", allow_module_level=True)

import rtspatial_amd as R  # noqa: E402
from rtspatial_amd import _native as N  # noqa: E402
from gpu_util import collect, dev, log  # noqa: E402

LEAF = 0x80000000


def _d2h(ptr, nbytes):
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipDeviceSynchronize()
    out = np.empty(nbytes, np.uint8)
    assert hip.hipMemcpy(out.ctypes.data, ptr, nbytes, 2) == 0  # hipMemcpyDeviceToHost
    return out


def _chain_boxes(dups):
    # Scene [0, 1]^2 (corner boxes at (0, 0) and (1, 1)); a centre
    # (x + 0.5) / 32767 quantizes to x, so the key is the interleave of the
    # chosen grid coordinates: one-bit keys 2^b, then `dups` boxes at 0.
    cs = []
    for b in range(30):
        gx, gy = (1 << (b // 2), 0) if b % 2 == 0 else (0, 1 << (b // 2))
        cs.append(((gx + 0.5) / 32767.0, (gy + 0.5) / 32767.0))
    cs += [(0.0, 0.0)] * dups + [(1.0, 1.0)]
    c = np.float32(cs)
    return np.concatenate([c, c], 1).astype(np.float32)


def test_deepest_tree_depth_and_lane_queries(oracle):
    log("deep tree: one-bit key chain over 2^17 equal keys")
    boxes = _chain_boxes(1 << 17)
    ix = R.SpatialIndex()
    ix.Init(R.Config(leaf_size=1, traversal=N.RTS_TRAVERSAL_THREAD))
    ix.Insert(dev(boxes))
    v = N.rts_tree_view()
    N.check(N.lib().rts_index_tree_view(ix._handle(), None, ctypes.byref(v)), "tree view")
    n = int(v.n)
    assert n == boxes.shape[0]
    raw = _d2h(v.nodes, 64 * (n - 1)).view(np.uint32).reshape(n - 1, 16)
    meta = raw[:, 8:12]  # (left, right, first, last) after the two child boxes
    depth = np.zeros(n - 1, np.int64)
    deepest = 0
    stack = [0]
    while stack:
        u = stack.pop()
        for c in (int(meta[u, 0]), int(meta[u, 1])):
            if c & LEAF:
                deepest = max(deepest, depth[u] + 1)
            else:
                depth[c] = depth[u] + 1
                stack.append(c)
    log(f"deep tree: {n} leaves, deepest leaf under {deepest} internal nodes")
    assert 40 <= deepest <= 60
    pts = np.unique(boxes[:, :2], axis=0)
    got, cnt = collect(ix, R.Predicate.kContains, pts)
    assert np.array_equal(got, oracle.bf_pairs(oracle.CONTAINS, boxes, pts))
    ix.Sync()  # raises if a walk flagged its stack or step bound
