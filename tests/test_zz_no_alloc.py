"""rts_index_query_static, the allocation-free query entry point, called
through the C ABI: its results equal those of the allocating entry points
(rts_index_query_points / _envelopes / _lines) on the same index and
queries, on two calls in a row with the second smaller (the shared static
scratch keeps the first call's contents), for every traversal kind the
plans pick (per-lane points, two-pass wave intersects / within, staged
segment walk) and both handlers; and its limits are refused before anything
is queued."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():
    pytest.skip("no GPU", allow_module_level=True)

import rtspatial_amd as R  # noqa: E402
from rtspatial_amd import _native as N  # noqa: E402
from rtspatial_amd import workloads as W  # noqa: E402
from gpu_util import canon, dev, log  # noqa: E402

P = R.Predicate
KIND_POINTS, KIND_ENVELOPES, KIND_LINES = N.RTS_QUERY_POINTS, N.RTS_QUERY_ENVELOPES, N.RTS_QUERY_LINES


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _call(ix, fn_name, pred, kind, q, handler, arg):
    L = N.lib()
    h = ix._handle()
    if fn_name == "static":
        return L.rts_index_query_static(h, int(pred), kind, q.data_ptr(), q.shape[0], handler, arg,
                                        _stream())
    fn = {KIND_POINTS: L.rts_index_query_points, KIND_ENVELOPES: L.rts_index_query_envelopes,
          KIND_LINES: L.rts_index_query_lines}[kind]
    return fn(h, int(pred), q.data_ptr(), q.shape[0], handler, arg, _stream())


def _count(ix, fn_name, pred, kind, q):
    c = R.SharedValue()
    c.set(0)
    N.check(_call(ix, fn_name, pred, kind, q, N.RTS_HANDLER_COUNT, c.data()), fn_name)
    return c.get()


def _pairs(ix, fn_name, pred, kind, q, cap):
    queue = R.Queue(cap)
    N.check(_call(ix, fn_name, pred, kind, q, N.RTS_HANDLER_COLLECT, queue.DeviceObject()), fn_name)
    n = queue.size()
    assert n <= cap
    return canon(queue.pairs()), n


@pytest.fixture(scope="module")
def index():
    rng = np.random.default_rng(4242)
    means = W.mixture_means(rng)
    boxes = W.mixture_boxes(rng, means, 400_000, np.log(1e-3))
    ix = R.SpatialIndex()
    ix.Init(R.Config())
    ix.Insert(dev(boxes))
    ix.Build()
    ix.Sync()
    return ix, boxes, rng, means


CASES = [("points", P.kContains, KIND_POINTS),
         ("intersects", P.kIntersects, KIND_ENVELOPES),
         ("within", P.kWithin, KIND_ENVELOPES),
         ("contains", P.kContains, KIND_ENVELOPES),
         ("segments", P.kIntersects, KIND_LINES)]


@pytest.mark.parametrize("name,pred,kind", CASES, ids=[c[0] for c in CASES])
def test_static_equals_allocating_two_calls(oracle, index, name, pred, kind):
    ix, boxes, rng, means = index
    # 300k then 70k queries: the second call is smaller, so the scratch still
    # holds the first call's counts, offsets and order beyond its end.
    for n in (300_000, 70_000):
        q = W.mixture_boxes(rng, means, n, np.log(4e-3))
        if kind == KIND_POINTS:
            q = np.ascontiguousarray(q[:, :2])
        dq = dev(q)
        log(f"static query {name}: {n} queries")
        c_alloc = _count(ix, "alloc", pred, kind, dq)
        c_static = _count(ix, "static", pred, kind, dq)
        assert c_static == c_alloc, (name, n)
        got_a, na = _pairs(ix, "alloc", pred, kind, dq, c_alloc + 16)
        got_s, ns = _pairs(ix, "static", pred, kind, dq, c_alloc + 16)
        assert na == ns == c_alloc
        assert np.array_equal(got_s, got_a), (name, n)
        if kind == KIND_LINES:
            want = oracle.bf_pairs(int(pred), boxes, q[:2000], lines=True)
            sel = got_s[got_s[:, 0] < 2000]
            assert np.array_equal(sel, want)
        else:
            want = oracle.RTree(boxes).pairs(int(pred), q)
            assert np.array_equal(got_s, want), (name, n)
    ix.Sync()


def test_static_limits_refused(index):
    ix, boxes, rng, means = index
    L = N.lib()
    q = dev(W.uniform_boxes(rng, (1 << 20) + 1, 0.01))
    c = R.SharedValue()
    rc = L.rts_index_query_static(ix._handle(), int(P.kIntersects), KIND_ENVELOPES, q.data_ptr(),
                                  q.shape[0], N.RTS_HANDLER_COUNT, c.data(), _stream())
    assert rc == N.RTS_ERR_INVALID
    assert b"1048576" in L.rts_last_error()
    # A pending mutation must be built by an allocating call first.
    small = R.SpatialIndex()
    small.Init(R.Config())
    small.Insert(dev(boxes[:1000]))
    rc = L.rts_index_query_static(small._handle(), int(P.kIntersects), KIND_ENVELOPES, q.data_ptr(),
                                  100, N.RTS_HANDLER_COUNT, c.data(), _stream())
    assert rc == N.RTS_ERR_INVALID
    assert b"rts_index_build" in L.rts_last_error()
    small.Build()
    N.check(L.rts_index_query_static(small._handle(), int(P.kIntersects), KIND_ENVELOPES,
                                     q.data_ptr(), 100, N.RTS_HANDLER_COUNT, c.data(), _stream()),
            "static")
    small.Sync()
