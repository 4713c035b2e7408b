"""Steady-state launches from a captured hipGraph: rts_index_query_static
(the allocation-free query entry point, which launches only kernels -- no
allocation, no memset or copy nodes) is captured once on a stream, and the
graph is replayed several times, with the program's own eager work and
synchronisations between replays; every replay's results must equal the
eager call's, which equal the oracle's."""
import ctypes
import ctypes.util

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
CAPTURE_THREAD_LOCAL = 1


def _hip():
    for name in ("libamdhip64.so", ctypes.util.find_library("amdhip64") or ""):
        if name:
            try:
                return ctypes.CDLL(name)
            except OSError:
                continue
    pytest.skip("libamdhip64 not loadable")


def _check(rc, what):
    assert rc == 0, f"{what} returned hipError {rc}"


@pytest.mark.parametrize("name,pred,kind,side", [
    ("intersects", P.kIntersects, N.RTS_QUERY_ENVELOPES, 4e-3),
    ("points", P.kContains, N.RTS_QUERY_POINTS, 0.0)])
def test_static_query_graph_replays(oracle, name, pred, kind, side):
    hip = _hip()
    R.profile_enable(False)  # no phase events in the captured sequence
    rng = np.random.default_rng(99)
    means = W.mixture_means(rng)
    boxes = W.mixture_boxes(rng, means, 300_000, np.log(1e-3))
    q = W.mixture_boxes(rng, means, 150_000, np.log(side) if side else np.log(1e-3))
    if kind == N.RTS_QUERY_POINTS:
        q = np.ascontiguousarray(q[:, :2])
    want = oracle.RTree(boxes).pairs(int(pred), q)
    log(f"graph replay {name}: {q.shape[0]} queries, {want.shape[0]} pairs")

    ix = R.SpatialIndex()
    ix.Init(R.Config())
    ix.Insert(dev(boxes))
    ix.Build()
    ix.Sync()
    dq = dev(q)
    L = N.lib()
    s = torch.cuda.Stream()
    h = ctypes.c_void_p(s.cuda_stream)
    counter = R.SharedValue()
    queue = R.Queue(want.shape[0] + 64)

    def calls():
        N.check(L.rts_index_query_static(ix._handle(), int(pred), kind, dq.data_ptr(), dq.shape[0],
                                         N.RTS_HANDLER_COUNT, counter.data(), h), "static count")
        N.check(L.rts_index_query_static(ix._handle(), int(pred), kind, dq.data_ptr(), dq.shape[0],
                                         N.RTS_HANDLER_COLLECT, queue.DeviceObject(), h),
                "static collect")

    def reset():
        with torch.cuda.stream(s):
            counter.tensor.zero_()
            queue.counter.zero_()
        s.synchronize()

    # Eager run on the stream (it also takes the scratch array's address).
    reset()
    calls()
    s.synchronize()
    assert counter.get() == want.shape[0]
    assert queue.size() == want.shape[0]
    assert np.array_equal(canon(queue.pairs()), want)

    graph = ctypes.c_void_p()
    exe = ctypes.c_void_p()
    _check(hip.hipStreamBeginCapture(h, CAPTURE_THREAD_LOCAL), "hipStreamBeginCapture")
    try:
        calls()
    finally:
        _check(hip.hipStreamEndCapture(h, ctypes.byref(graph)), "hipStreamEndCapture")
    _check(hip.hipGraphInstantiate(ctypes.byref(exe), graph, None, None, ctypes.c_size_t(0)),
           "hipGraphInstantiate")
    try:
        for rep in range(3):
            reset()
            _check(hip.hipGraphLaunch(exe, h), "hipGraphLaunch")
            s.synchronize()
            log(f"graph replay {name}: replay {rep} done")
            assert counter.get() == want.shape[0], rep
            assert queue.size() == want.shape[0], rep
            assert np.array_equal(canon(queue.pairs()), want), rep
            # Eager work of the program's own between replays.
            reset()
            calls()
            s.synchronize()
            assert counter.get() == want.shape[0]
    finally:
        hip.hipGraphExecDestroy(exe)
        hip.hipGraphDestroy(graph)
    ix.Sync()
