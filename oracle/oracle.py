"""ctypes wrapper over liboracle.so, the CPU oracle.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the rtspatial_amd package.

Array conventions (numpy):
  boxes    float32 [n, 4]  (x0, y0, x1, y1)  -- Envelope<Point<float,2>>,
           include/rtspatial/geom/envelope.cuh:13-109 (min corner then max corner)
  points   float32 [n, 2]  (x, y)            -- Point<float,2>, geom/point.cuh:11-42
  segments float32 [n, 4]  (x1, y1, x2, y2)  -- Line<Point<float,2>>, geom/line.cuh:10-27;
           pass lines=True (INTERSECTS only: CONTAINS and WITHIN match nothing)
  alive    uint8   [n] or None (deleted boxes are 0)
  pairs    uint32  [k, 2]  (query_id, box_id), sorted by (query_id, box_id)

Double-coordinate indexes (SpatialIndex<double, 2>) are checked with the
*64 functions: float64 arrays, exact double predicates.

Predicates: INTERSECTS = 0, CONTAINS = 1 (data box contains the query,
spatial_index.cuh:414-441 with shaders_contains_envelope_query_2d.cu:22),
WITHIN = 2 (data box inside the query envelope).
"""
import ctypes
import os

import numpy as np

INTERSECTS, CONTAINS, WITHIN = 0, 1, 2
POINT, ENVELOPE, LINE = 0, 1, 2

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: run `make -C oracle` (or __graft_entry__.build())")
        L = ctypes.CDLL(path)
        u64, i32, f32, vp = ctypes.c_uint64, ctypes.c_int, ctypes.c_float, ctypes.c_void_p
        L.orc_ref_uniform_boxes_f32.argtypes = [u64, f32, f32, vp]
        L.orc_ref_uniform_points_f32.argtypes = [u64, vp]
        L.orc_ref_update_ids.argtypes = [u64, u64, vp]
        L.orc_eval.argtypes = [i32, i32, vp, vp]
        L.orc_eval.restype = i32
        L.orc_splitmix64.argtypes = [u64]
        L.orc_splitmix64.restype = u64
        L.orc_bf_summary.argtypes = [i32, i32, vp, u64, vp, vp, u64, i32, vp, vp, vp]
        L.orc_bf_summary.restype = u64
        L.orc_bf_pairs.argtypes = [i32, i32, vp, u64, vp, vp, u64, i32, vp, u64]
        L.orc_bf_pairs.restype = u64
        L.orc_bf_summary_f64.argtypes = [i32, i32, vp, u64, vp, vp, u64, i32, vp, vp, vp]
        L.orc_bf_summary_f64.restype = u64
        L.orc_bf_pairs_f64.argtypes = [i32, i32, vp, u64, vp, vp, u64, i32, vp, u64]
        L.orc_bf_pairs_f64.restype = u64
        L.orc_eval_f64.argtypes = [i32, i32, vp, vp]
        L.orc_eval_f64.restype = i32
        L.orc_rtree_build.argtypes = [vp, u64, vp, i32, i32]
        L.orc_rtree_build.restype = vp
        L.orc_rtree_update.argtypes = [vp, vp, vp, u64]
        L.orc_rtree_update.restype = u64
        L.orc_rtree_checksums.argtypes = [vp, i32, i32, vp, u64, i32, vp, vp]
        L.orc_rtree_checksums.restype = u64
        L.orc_ghash.argtypes = [u64]
        L.orc_ghash.restype = u64
        L.orc_qmult.argtypes = [u64]
        L.orc_qmult.restype = u64
        L.orc_rtree_free.argtypes = [vp]
        L.orc_rtree_summary.argtypes = [vp, i32, i32, vp, u64, i32, vp, vp, vp]
        L.orc_rtree_summary.restype = u64
        L.orc_rtree_pairs.argtypes = [vp, i32, i32, vp, u64, i32, vp, u64]
        L.orc_rtree_pairs.restype = u64
        L.orc_refray_intersects.argtypes = [vp, u64, vp, u64, i32, vp, u64, vp, vp, u64, vp]
        L.orc_refray_intersects.restype = u64
        _LIB = L
    return _LIB


def _p(a):
    return None if a is None else a.ctypes.data


def _f32(a, cols):
    a = np.ascontiguousarray(a, dtype=np.float32)
    assert a.ndim == 2 and a.shape[1] == cols, a.shape
    return a


def _kind(queries, lines=False):
    if lines:
        assert queries.shape[1] == 4, queries.shape
        return LINE
    return POINT if queries.shape[1] == 2 else ENVELOPE


def _alive(alive, n):
    if alive is None:
        return None
    a = np.ascontiguousarray(alive, dtype=np.uint8)
    assert a.shape == (n,)
    return a


# ------------------------------------------------------------- generators --
def ref_uniform_boxes(n, max_w, max_h):
    """test/test_commons.h:22-49, GenerateUniformBoxes<float>(n, max_w, max_h)."""
    out = np.empty((n, 4), np.float32)
    lib().orc_ref_uniform_boxes_f32(n, max_w, max_h, _p(out))
    return out


def ref_uniform_points(n):
    """test/test_commons.h:51-67, GenerateUniformPoints<float>(n)."""
    out = np.empty((n, 2), np.float32)
    lib().orc_ref_uniform_points_f32(n, _p(out))
    return out


def ref_update_ids(n1, n_updates):
    """test/test_envelope_queries.h:121-134, the distinct ids drawn for Update."""
    out = np.empty(n_updates, np.uint64)
    lib().orc_ref_update_ids(n1, n_updates, _p(out))
    return out


def splitmix64(x):
    return lib().orc_splitmix64(int(x))


def evaluate(pred, box, query, lines=False):
    b = np.ascontiguousarray(box, np.float32)
    q = np.ascontiguousarray(query, np.float32)
    kind = LINE if lines else (POINT if q.size == 2 else ENVELOPE)
    return bool(lib().orc_eval(pred, kind, _p(b), _p(q)))


# ------------------------------------------------------------ brute force --
def bf_summary(pred, boxes, queries, alive=None, threads=0, want_hash=True, lines=False):
    """Total count, per-query counts and per-query (sum, xor) of
    splitmix64((query_id << 32) | box_id) over the matching boxes."""
    boxes = _f32(boxes, 4)
    queries = np.ascontiguousarray(queries, dtype=np.float32)
    nq = queries.shape[0]
    per_q = np.zeros(nq, np.uint64)
    hs = np.zeros(nq, np.uint64) if want_hash else None
    hx = np.zeros(nq, np.uint64) if want_hash else None
    al = _alive(alive, boxes.shape[0])
    total = lib().orc_bf_summary(pred, _kind(queries, lines), _p(boxes), boxes.shape[0], _p(al),
                                 _p(queries), nq, threads, _p(per_q), _p(hs), _p(hx))
    return int(total), per_q, hs, hx


def bf_count(pred, boxes, queries, alive=None, threads=0, lines=False):
    boxes = _f32(boxes, 4)
    queries = np.ascontiguousarray(queries, dtype=np.float32)
    al = _alive(alive, boxes.shape[0])
    return int(lib().orc_bf_summary(pred, _kind(queries, lines), _p(boxes), boxes.shape[0], _p(al),
                                    _p(queries), queries.shape[0], threads, None, None, None))


def bf_pairs(pred, boxes, queries, alive=None, threads=0, lines=False):
    boxes = _f32(boxes, 4)
    queries = np.ascontiguousarray(queries, dtype=np.float32)
    al = _alive(alive, boxes.shape[0])
    total = bf_count(pred, boxes, queries, alive, threads, lines)
    out = np.empty((max(total, 1), 2), np.uint32)
    got = lib().orc_bf_pairs(pred, _kind(queries, lines), _p(boxes), boxes.shape[0], _p(al),
                             _p(queries), queries.shape[0], threads, _p(out), total)
    assert got == total
    return out[:total]


# ------------------------------------------------- double coordinates --
# SpatialIndex<double, 2>: boxes float64 [n, 4], points float64 [n, 2], the
# predicates evaluated exactly in double precision.
def _f64(a, cols):
    a = np.ascontiguousarray(a, dtype=np.float64)
    assert a.ndim == 2 and a.shape[1] == cols, a.shape
    return a


def evaluate64(pred, box, query, lines=False):
    b = np.ascontiguousarray(box, np.float64)
    q = np.ascontiguousarray(query, np.float64)
    kind = LINE if lines else (POINT if q.size == 2 else ENVELOPE)
    return bool(lib().orc_eval_f64(pred, kind, _p(b), _p(q)))


def bf_summary64(pred, boxes, queries, alive=None, threads=0, want_hash=True, lines=False):
    boxes = _f64(boxes, 4)
    queries = np.ascontiguousarray(queries, dtype=np.float64)
    nq = queries.shape[0]
    per_q = np.zeros(nq, np.uint64)
    hs = np.zeros(nq, np.uint64) if want_hash else None
    hx = np.zeros(nq, np.uint64) if want_hash else None
    al = _alive(alive, boxes.shape[0])
    total = lib().orc_bf_summary_f64(pred, _kind(queries, lines), _p(boxes), boxes.shape[0], _p(al),
                                     _p(queries), nq, threads, _p(per_q), _p(hs), _p(hx))
    return int(total), per_q, hs, hx


def bf_count64(pred, boxes, queries, alive=None, threads=0, lines=False):
    return bf_summary64(pred, boxes, queries, alive, threads, want_hash=False, lines=lines)[0]


def bf_pairs64(pred, boxes, queries, alive=None, threads=0, lines=False):
    """(query_id, box_id) pairs sorted by (query, box), exact double predicates."""
    boxes = _f64(boxes, 4)
    queries = np.ascontiguousarray(queries, dtype=np.float64)
    al = _alive(alive, boxes.shape[0])
    total = bf_count64(pred, boxes, queries, alive, threads, lines)
    out = np.empty((max(total, 1), 2), np.uint32)
    got = lib().orc_bf_pairs_f64(pred, _kind(queries, lines), _p(boxes), boxes.shape[0], _p(al),
                                 _p(queries), queries.shape[0], threads, _p(out), total)
    assert got == total
    return out[:total]


# ------------------------------------------------------- line segments --
# Line<Point<float,2>> queries (geom/line.cuh:10-27; the reference's line
# intersection program reports nothing, shaders_intersects_line_query_2d.cu:
# 14-41). A closed segment (x1, y1, x2, y2) meets a closed box iff their
# bounding boxes overlap and the box's corners are not all strictly on one
# side of the segment's line; evaluated here in exact rational arithmetic
# (finite coordinates; a segment with a NaN coordinate meets nothing).
def segment_meets_box(seg, box):
    from fractions import Fraction as F
    if any(v != v for v in (float(c) for c in seg)):
        return False  # a segment with a NaN coordinate meets nothing
    x1, y1, x2, y2 = (F(float(v)) for v in seg)
    bx0, by0, bx1, by1 = (float(v) for v in box)
    if not (min(x1, x2) <= bx1 and max(x1, x2) >= bx0 and min(y1, y2) <= by1
            and max(y1, y2) >= by0):
        return False  # also for a NaN box: every comparison is false
    dx, dy = x2 - x1, y2 - y1
    sides = set()
    for cx, cy in ((bx0, by0), (bx1, by0), (bx1, by1), (bx0, by1)):
        o = dx * (F(cy) - y1) - dy * (F(cx) - x1)
        sides.add((o > 0) - (o < 0))
    return sides != {1} and sides != {-1}


def segment_pairs(boxes, segs, alive=None):
    """(query_id, box_id) pairs, sorted, of segments meeting boxes: numpy
    bounding-box prefilter, exact test on the candidates (small inputs)."""
    b = np.asarray(boxes)
    sg = np.asarray(segs)
    ok = np.ones(b.shape[0], bool) if alive is None else np.asarray(alive).astype(bool)
    out = []
    for qi, seg in enumerate(sg):
        sx0, sx1 = min(seg[0], seg[2]), max(seg[0], seg[2])
        sy0, sy1 = min(seg[1], seg[3]), max(seg[1], seg[3])
        cand = np.nonzero(ok & (b[:, 0] <= sx1) & (b[:, 2] >= sx0) & (b[:, 1] <= sy1)
                          & (b[:, 3] >= sy0))[0]
        for g in cand:
            if segment_meets_box(seg, b[g]):
                out.append((qi, g))
    return np.asarray(out, dtype=np.uint32).reshape(-1, 2)


# ----------------------------------------------------------------- R-tree --
class RTree:
    """STR bulk-loaded R-tree (fan-out 16 by default) with exact FP32 bounds.
    `threads` (0: every hardware thread) is used by the build and refits."""

    def __init__(self, boxes, alive=None, fanout=16, threads=0):
        self._boxes = _f32(boxes, 4)
        al = _alive(alive, self._boxes.shape[0])
        self._h = lib().orc_rtree_build(_p(self._boxes), self._boxes.shape[0], _p(al), fanout,
                                        threads)

    def __del__(self):
        if getattr(self, "_h", None):
            lib().orc_rtree_free(self._h)
            self._h = None

    def summary(self, pred, queries, threads=0, want_hash=True):
        queries = np.ascontiguousarray(queries, dtype=np.float32)
        nq = queries.shape[0]
        per_q = np.zeros(nq, np.uint64)
        hs = np.zeros(nq, np.uint64) if want_hash else None
        hx = np.zeros(nq, np.uint64) if want_hash else None
        total = lib().orc_rtree_summary(self._h, pred, _kind(queries), _p(queries), nq, threads,
                                        _p(per_q), _p(hs), _p(hx))
        return int(total), per_q, hs, hx

    def count(self, pred, queries, threads=0):
        queries = np.ascontiguousarray(queries, dtype=np.float32)
        return int(lib().orc_rtree_summary(self._h, pred, _kind(queries), _p(queries),
                                           queries.shape[0], threads, None, None, None))

    def checksums(self, pred, queries, threads=0):
        """(total, per-query counts, per-query ghash sums): the factored
        checksum of each query's result set (see pair_checksums); envelope
        intersects / within count subtrees inside the query whole."""
        queries = np.ascontiguousarray(queries, dtype=np.float32)
        nq = queries.shape[0]
        cnt = np.zeros(nq, np.uint64)
        gs = np.zeros(nq, np.uint64)
        total = lib().orc_rtree_checksums(self._h, pred, _kind(queries), _p(queries), nq, threads,
                                          _p(cnt), _p(gs))
        return int(total), cnt, gs

    def update(self, ids, boxes):
        """Boxes `ids` take the new envelopes (NaN: deleted), then the tree is
        refitted. Ids not in the tree are ignored; returns how many applied."""
        ids = np.ascontiguousarray(ids, dtype=np.uint64)
        b = _f32(boxes, 4)
        assert b.shape[0] == ids.shape[0]
        return int(lib().orc_rtree_update(self._h, _p(ids), _p(b), ids.shape[0]))

    def delete(self, ids):
        ids = np.ascontiguousarray(ids, dtype=np.uint64)
        return self.update(ids, np.full((ids.shape[0], 4), np.nan, np.float32))

    def pairs(self, pred, queries, threads=0):
        queries = np.ascontiguousarray(queries, dtype=np.float32)
        total = self.count(pred, queries, threads)
        out = np.empty((max(total, 1), 2), np.uint32)
        got = lib().orc_rtree_pairs(self._h, pred, _kind(queries), _p(queries), queries.shape[0],
                                    threads, _p(out), total)
        assert got == total
        return out[:total]


# ------------------------------------------------ reference ray formulation --
def refray_intersects(boxes, queries, threads=0, cap=1024):
    """The reference's two-pass ray formulation of envelope intersects
    (ray_params.h:53-167, shaders_intersects_envelope_query_2d.cu:14-128).
    Returns (count, extra_pairs, missing_pairs) relative to the exact predicate."""
    boxes = _f32(boxes, 4)
    queries = _f32(queries, 4)
    extra = np.zeros((cap, 2), np.uint32)
    missing = np.zeros((cap, 2), np.uint32)
    ne, nm = ctypes.c_uint64(0), ctypes.c_uint64(0)
    total = lib().orc_refray_intersects(_p(boxes), boxes.shape[0], _p(queries), queries.shape[0],
                                        threads, _p(extra), cap, ctypes.byref(ne), _p(missing),
                                        cap, ctypes.byref(nm))
    return int(total), extra[: min(ne.value, cap)], missing[: min(nm.value, cap)]


# ---------------------------------------------------- summaries of pair sets --
def summarize_pairs(pairs, nq):
    """Per-query count and (sum, xor) hash summaries of a (query, box) pair array."""
    pairs = np.asarray(pairs, dtype=np.uint64).reshape(-1, 2)
    q = pairs[:, 0]
    keys = (q << np.uint64(32)) | pairs[:, 1]
    h = _splitmix_np(keys)
    cnt = np.bincount(q.astype(np.int64), minlength=nq).astype(np.uint64)
    hs = np.zeros(nq, np.uint64)
    np.add.at(hs, q.astype(np.int64), h)
    hx = np.zeros(nq, np.uint64)
    np.bitwise_xor.at(hx, q.astype(np.int64), h)
    return cnt, hs, hx


def _splitmix_np(x):
    x = np.asarray(x, dtype=np.uint64)
    with np.errstate(over="ignore"):
        x = x + np.uint64(0x9E3779B97F4A7C15)
        x = (x ^ (x >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        x = (x ^ (x >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return x ^ (x >> np.uint64(31))


# ------------------------------------------------ factored pair checksum --
# A pair (query q, box b) hashes to qmult(q) * ghash(b) mod 2^64 with
# ghash(b) = splitmix64(b ^ 0xD1B54A32D192ED03) and qmult(q) = splitmix64(q) | 1.
# A query's result set is summarised by its count and gsum = sum of ghash(b);
# a request's checksum is sum over its queries of qmult(q) * gsum(q).
_GSALT = np.uint64(0xD1B54A32D192ED03)


def ghash_np(b):
    return _splitmix_np(np.asarray(b, dtype=np.uint64) ^ _GSALT)


def qmult_np(q):
    return _splitmix_np(np.asarray(q, dtype=np.uint64)) | np.uint64(1)


def pair_checksums(pairs, nq):
    """Per-query counts and ghash sums of a (query_id, box_id) pair array."""
    pairs = np.asarray(pairs, dtype=np.uint64).reshape(-1, 2)
    q = pairs[:, 0].astype(np.int64)
    cnt = np.bincount(q, minlength=nq).astype(np.uint64)
    gs = np.zeros(nq, np.uint64)
    np.add.at(gs, q, ghash_np(pairs[:, 1]))
    return cnt, gs


def request_checksum(gsums, qids=None):
    """sum over queries of qmult(q) * gsum(q), q = 0..n-1 or `qids`."""
    gsums = np.asarray(gsums, dtype=np.uint64)
    q = np.arange(gsums.shape[0], dtype=np.uint64) if qids is None else np.asarray(qids, np.uint64)
    with np.errstate(over="ignore"):
        return int(np.sum(qmult_np(q) * gsums, dtype=np.uint64))
