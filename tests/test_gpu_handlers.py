"""Result handlers compiled by the caller (include/rtspatial/device/query.h),
through the example library rtspatial_amd/examples/librts_user_handlers.so,
against the oracle. The reference's equivalent is its handler plugin:
rtspatial_handle_* functions compiled into its shaders
(default_handlers/*.h, test/CMakeLists.txt:8).

  Summary   per query: count, sum of splitmix64((q << 32) | g), min geom id
  CountDone per query: count, from the per-query hook alone
  reference counting functions: one counter for all results
"""
import ctypes
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():
    pytest.skip("no GPU", allow_module_level=True)

import rtspatial_amd as R  # noqa: E402
from gpu_util import dev, log, make_index  # noqa: E402

N = R._native
P = R.Predicate
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "rtspatial_amd", "examples", "librts_user_handlers.so")
THREAD, WAVE = N.RTS_TRAVERSAL_THREAD, N.RTS_TRAVERSAL_WAVE


@pytest.fixture(scope="module")
def X():
    assert os.path.exists(LIB), "build first: python -c 'import __graft_entry__ as g; g.build()'"
    N.lib()  # the example links the library; load it first from the package
    L = ctypes.CDLL(LIB)
    vp, u64, i32, u32 = ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int, ctypes.c_uint32
    L.rtsx_summary.argtypes = [vp, i32, vp, u64, i32, i32, u32, vp, vp, vp, vp]
    L.rtsx_count.argtypes = [vp, i32, vp, u64, i32, i32, u32, vp, vp]
    L.rtsx_reference_count.argtypes = [vp, i32, vp, u64, i32, i32, vp, vp]
    L.rtsx_last_error.restype = ctypes.c_char_p
    return L


def run_summary(X, ix, pred, q, trav):
    n = q.shape[0]
    count = torch.zeros(n, dtype=torch.int64, device="cuda")
    hsum = torch.zeros(n, dtype=torch.int64, device="cuda")
    mn = torch.full((n,), -1, dtype=torch.int32, device="cuda")
    rc = X.rtsx_summary(ix._handle(), int(pred), q.data_ptr(), n, 1 if q.shape[1] == 2 else 0,
                        1 if q.dtype == torch.float64 else 0, trav, count.data_ptr(),
                        hsum.data_ptr(), mn.data_ptr(), None)
    assert rc == 0, X.rtsx_last_error()
    done = torch.full((n,), -7, dtype=torch.int64, device="cuda")
    rc = X.rtsx_count(ix._handle(), int(pred), q.data_ptr(), n, 1 if q.shape[1] == 2 else 0,
                      1 if q.dtype == torch.float64 else 0, trav, done.data_ptr(), None)
    assert rc == 0, X.rtsx_last_error()
    ix.Sync()
    u64 = lambda t: t.cpu().numpy().view(np.uint64)
    return u64(count), u64(hsum), mn.cpu().numpy().view(np.uint32), u64(done)


def oracle_summary(oracle, pred, b, q, alive=None):
    if b.dtype == np.float64:
        tot, cnt, hs, _ = oracle.bf_summary64(int(pred), b, q, alive=alive)
        pairs = oracle.bf_pairs64(int(pred), b, q, alive=alive)
    else:
        tot, cnt, hs, _ = oracle.bf_summary(int(pred), b, q, alive=alive)
        pairs = oracle.bf_pairs(int(pred), b, q, alive=alive)
    mn = np.full(q.shape[0], 0xFFFFFFFF, np.uint32)
    if pairs.shape[0]:
        first = np.r_[True, pairs[1:, 0] != pairs[:-1, 0]]  # pairs sorted by (query, box)
        mn[pairs[first, 0]] = pairs[first, 1]
    return tot, cnt, hs, mn


def inputs(dtype, seed=50):
    from rtspatial_amd import workloads as W
    rng = np.random.default_rng(seed)
    means = W.mixture_means(rng)
    b = W.mixture_boxes(rng, means, 80_000, np.log(2e-3), dtype=dtype)
    q = W.mixture_boxes(rng, means, 3_000, np.log(1e-2), dtype=dtype)
    return b, q, q[:, :2].copy()


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("trav", [THREAD, WAVE])
def test_handlers_match_oracle(X, oracle, dtype, trav):
    b, q, pts = inputs(dtype)
    log(f"user handlers dtype={np.dtype(dtype).name} traversal={trav}")
    # Morton-ordered queries (the order comes from rts_index_order_queries) with
    # the wave traversal, input order with the lane one.
    ix = make_index(b, leaf_size=8, sort_queries=1 if trav == WAVE else 2)
    for pred, qq in ((P.kIntersects, q), (P.kContains, q), (P.kWithin, q), (P.kContains, pts)):
        tot, cnt, hs, mn = oracle_summary(oracle, pred, b, qq)
        c, h, m, d = run_summary(X, ix, pred, dev(qq), trav)
        assert np.array_equal(c, cnt) and np.array_equal(h, hs), (int(pred), qq.shape)
        assert np.array_equal(m, mn) and np.array_equal(d, cnt), (int(pred), qq.shape)
        counter = torch.zeros(1, dtype=torch.int64, device="cuda")
        if pred != P.kWithin:
            rc = X.rtsx_reference_count(ix._handle(), int(pred), dev(qq).data_ptr(), qq.shape[0],
                                        1 if qq.shape[1] == 2 else 0,
                                        1 if dtype == np.float64 else 0, counter.data_ptr(), None)
            assert rc == 0, X.rtsx_last_error()
            assert int(counter.item()) == tot


def test_handlers_after_delete_update_and_rounding_mode(X, oracle):
    b, q, pts = inputs(np.float64, seed=51)
    log("user handlers after mutations, double index")
    rng = np.random.default_rng(52)
    ix = make_index(b)
    ids = rng.choice(b.shape[0], 4000, replace=False).astype(np.int64)
    new = b[rng.integers(0, b.shape[0], 4000)] + 1e-3
    ix.Update((dev(ids), dev(new)))
    dead = rng.choice(b.shape[0], 3000, replace=False).astype(np.int64)
    ix.Delete(dev(dead))
    b2 = b.copy()
    b2[ids] = new
    alive = np.ones(b.shape[0], np.uint8)
    alive[dead] = 0
    for trav in (THREAD, WAVE):
        _, cnt, hs, mn = oracle_summary(oracle, P.kIntersects, b2, q, alive)
        c, h, m, d = run_summary(X, ix, P.kIntersects, dev(q), trav)
        assert np.array_equal(c, cnt) and np.array_equal(h, hs) and np.array_equal(m, mn)
        assert np.array_equal(d, cnt)
    # an index that rounds doubles to floats gives the float oracle's answers
    rr = make_index(b, coord=N.RTS_COORD_F64_AS_F32)
    _, cnt, hs, mn = oracle_summary(oracle, P.kWithin, b.astype(np.float32), q.astype(np.float32))
    c, h, m, d = run_summary(X, rr, P.kWithin, dev(q), WAVE)
    assert np.array_equal(c, cnt) and np.array_equal(h, hs) and np.array_equal(m, mn)


def test_handler_errors_and_empty_index(X):
    log("user handler errors")
    ix = make_index(np.zeros((0, 4), np.float32))
    q = dev(np.random.default_rng(1).random((50, 2)).astype(np.float32))
    done = torch.full((50,), -7, dtype=torch.int64, device="cuda")
    assert X.rtsx_count(ix._handle(), int(P.kContains), q.data_ptr(), 50, 1, 0, 0,
                        done.data_ptr(), None) == 0
    torch.cuda.synchronize()
    assert int(done.abs().sum().item()) == 0  # every query reported, with 0 results
    cnt = torch.zeros(50, dtype=torch.int64, device="cuda")
    # points take only contains; double queries on a float index; within has no
    # reference handler function
    assert X.rtsx_count(ix._handle(), int(P.kIntersects), q.data_ptr(), 50, 1, 0, 0,
                        cnt.data_ptr(), None) == 1
    assert b"kContains" in X.rtsx_last_error()
    assert X.rtsx_count(ix._handle(), int(P.kContains), q.data_ptr(), 25, 1, 1, 0,
                        cnt.data_ptr(), None) == 1
    assert b"double queries on a float index" in X.rtsx_last_error()
    e = dev(np.random.default_rng(2).random((10, 4)).astype(np.float32))
    assert X.rtsx_reference_count(ix._handle(), int(P.kWithin), e.data_ptr(), 10, 0, 0,
                                  cnt.data_ptr(), None) == 1
    assert b"kWithin" in X.rtsx_last_error()
