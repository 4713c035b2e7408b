"""ctypes binding of librtspatial_amd.so (the C ABI declared in include/rts_c.h).

The library must have been built (python -m rtspatial_amd.build); there is no
CPU fallback: loading fails loudly when it is missing.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
# RTS_LIBRARY names another build of the same library (e.g. an instrumented one).
LIB_PATH = os.environ.get("RTS_LIBRARY") or os.path.join(_HERE, "librtspatial_amd.so")

RTS_OK = 0
RTS_ERR_INVALID, RTS_ERR_HIP, RTS_ERR_DEVICE, RTS_ERR_MISSING = 1, 2, 3, 4
RTS_INTERSECTS, RTS_CONTAINS, RTS_WITHIN = 0, 1, 2
RTS_HANDLER_COUNT, RTS_HANDLER_COLLECT = 0, 1
RTS_REBUILD_SAH, RTS_REBUILD_NEVER, RTS_REBUILD_ALWAYS = 0, 1, 2
RTS_TRAVERSAL_AUTO, RTS_TRAVERSAL_THREAD, RTS_TRAVERSAL_WAVE = 0, 1, 2
RTS_COORD_F32, RTS_COORD_F64, RTS_COORD_F64_AS_F32 = 0, 1, 2
RTS_QUERY_ENVELOPES, RTS_QUERY_POINTS, RTS_QUERY_LINES = 0, 1, 2


class rts_config(ctypes.Structure):
    _fields_ = [
        ("max_geometries", ctypes.c_uint64),
        ("max_queries", ctypes.c_uint64),
        ("preallocate", ctypes.c_int32),
        ("prefer_fast_build_geom", ctypes.c_int32),
        ("prefer_fast_build_query", ctypes.c_int32),
        ("compact", ctypes.c_int32),
        ("geom_sample_rate", ctypes.c_float),
        ("query_sample_rate", ctypes.c_float),
        ("intersect_cost_weight", ctypes.c_float),
        ("max_geom_samples", ctypes.c_uint32),
        ("max_query_samples", ctypes.c_uint32),
        ("max_parallelism", ctypes.c_uint32),
        ("leaf_size", ctypes.c_uint32),
        ("rebuild_policy", ctypes.c_uint32),
        ("sah_rebuild_ratio", ctypes.c_float),
        ("traversal", ctypes.c_uint32),
        ("sort_queries", ctypes.c_uint32),
        ("coord", ctypes.c_uint32),
        ("insert_delta_pct", ctypes.c_uint32),
    ]


class rts_tree_view(ctypes.Structure):
    _fields_ = [
        ("leaf_boxes", ctypes.c_void_p),
        ("leaf_ids", ctypes.c_void_p),
        ("nodes", ctypes.c_void_p),
        ("root_box", ctypes.c_void_p),
        ("root_live", ctypes.c_void_p),
        ("error", ctypes.c_void_p),
        ("boxes64", ctypes.c_void_p),
        ("n", ctypes.c_uint32),
        ("leaf_size", ctypes.c_uint32),
        ("coord", ctypes.c_uint32),
        ("reserved", ctypes.c_uint32),
    ]


class rts_query_plan(ctypes.Structure):
    _fields_ = [
        ("traversal", ctypes.c_uint32),
        ("sort_queries", ctypes.c_uint32),
        ("leaf_size", ctypes.c_uint32),
        ("reserved", ctypes.c_uint32),
    ]


class rts_memory_usage(ctypes.Structure):
    _fields_ = [
        ("geometries", ctypes.c_uint64),
        ("aabbs", ctypes.c_uint64),
        ("bvh", ctypes.c_uint64),
        ("total", ctypes.c_uint64),
    ]


class rts_dev_queue(ctypes.Structure):
    _fields_ = [
        ("data", ctypes.c_void_p),
        ("capacity", ctypes.c_uint64),
        ("counter", ctypes.c_void_p),
    ]


class rts_tune_config(ctypes.Structure):
    _fields_ = [
        ("sample", ctypes.c_uint32),
        ("leaf_sizes", ctypes.c_uint32),
        ("reps", ctypes.c_uint32),
        ("apply", ctypes.c_int32),
    ]


class rts_tune_result(ctypes.Structure):
    _fields_ = [
        ("leaf_size", ctypes.c_uint32),
        ("traversal", ctypes.c_uint32),
        ("sort_queries", ctypes.c_uint32),
        ("plans", ctypes.c_uint32),
        ("sample", ctypes.c_uint64),
        ("sample_results", ctypes.c_uint64),
        ("best_ms", ctypes.c_double),
        ("default_ms", ctypes.c_double),
    ]


class rts_index_stats(ctypes.Structure):
    _fields_ = [
        ("n_boxes", ctypes.c_uint64),
        ("n_tree", ctypes.c_uint64),
        ("builds", ctypes.c_uint64),
        ("refits", ctypes.c_uint64),
        ("sah_at_build", ctypes.c_double),
        ("sah_current", ctypes.c_double),
        ("stale", ctypes.c_int32),
        ("delta_boxes", ctypes.c_uint32),
    ]


class rts_request(ctypes.Structure):
    _fields_ = [
        ("arrival_s", ctypes.c_double),
        ("queries", ctypes.c_void_p),
        ("n", ctypes.c_uint64),
        ("points", ctypes.c_int32),
        ("predicate", ctypes.c_int32),
        ("out", ctypes.c_void_p),
        ("out_capacity", ctypes.c_uint64),
    ]


class rts_update_batch(ctypes.Structure):
    _fields_ = [("arrival_s", ctypes.c_double), ("updates", ctypes.c_void_p), ("n", ctypes.c_uint64)]


class rts_request_result(ctypes.Structure):
    _fields_ = [
        ("admit_s", ctypes.c_double),
        ("done_s", ctypes.c_double),
        ("count", ctypes.c_uint64),
        ("hash", ctypes.c_uint64),
        ("epoch", ctypes.c_uint32),
        ("relaunches", ctypes.c_uint32),
        ("truncated", ctypes.c_int32),
        ("pad", ctypes.c_int32),
    ]


class rts_serve_config(ctypes.Structure):
    _fields_ = [
        ("mode", ctypes.c_uint32),
        ("slots", ctypes.c_uint32),
        ("max_batch_requests", ctypes.c_uint32),
        ("use_graphs", ctypes.c_uint32),
        ("max_batch_queries", ctypes.c_uint64),
        ("chunk_queries", ctypes.c_uint64),
        ("arena_pairs", ctypes.c_uint64),
        ("hash_results", ctypes.c_uint32),
        ("traversal", ctypes.c_uint32),
        ("head_share_pct", ctypes.c_uint32),
        ("sort_min_tiles", ctypes.c_uint32),
    ]


class rts_serve_stats(ctypes.Structure):
    _fields_ = [
        ("wall_s", ctypes.c_double),
        ("requests", ctypes.c_uint64),
        ("queries", ctypes.c_uint64),
        ("results", ctypes.c_uint64),
        ("batches", ctypes.c_uint64),
        ("relaunches", ctypes.c_uint64),
        ("updates", ctypes.c_uint64),
    ]


RTS_SERVE_BATCHED, RTS_SERVE_PLAIN = 0, 1

RTS_PHASES = ["build_morton", "build_sort", "build_tree", "refit", "mutate", "query_order",
              "traverse", "query_count", "batch_count", "batch_write"]


class rts_profile(ctypes.Structure):
    _fields_ = [("count", ctypes.c_uint64 * 12), ("ms", ctypes.c_double * 12)]


# (name, restype, argtypes) of every exported entry point declared in rts_c.h.
_VP, _U64, _I32 = ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int
SIGNATURES = [
    ("rts_config_default", None, [ctypes.POINTER(rts_config)]),
    ("rts_index_create", _I32, [ctypes.POINTER(rts_config), ctypes.POINTER(_VP)]),
    ("rts_index_destroy", _I32, [_VP]),
    ("rts_index_clear", _I32, [_VP]),
    ("rts_index_insert", _I32, [_VP, _VP, _U64, _VP]),
    ("rts_index_delete", _I32, [_VP, _VP, _U64, _VP]),
    ("rts_index_update", _I32, [_VP, _VP, _U64, _VP]),
    ("rts_index_query_points", _I32, [_VP, _I32, _VP, _U64, _I32, _VP, _VP]),
    ("rts_index_query_envelopes", _I32, [_VP, _I32, _VP, _U64, _I32, _VP, _VP]),
    ("rts_pairs_checksum", _I32, [_VP, _U64, _U64, _VP, _VP, _VP]),
    ("rts_index_insert_d", _I32, [_VP, _VP, _U64, _VP]),
    ("rts_index_update_d", _I32, [_VP, _VP, _U64, _VP]),
    ("rts_index_query_points_d", _I32, [_VP, _I32, _VP, _U64, _I32, _VP, _VP]),
    ("rts_index_query_envelopes_d", _I32, [_VP, _I32, _VP, _U64, _I32, _VP, _VP]),
    ("rts_index_query_lines", _I32, [_VP, _I32, _VP, _U64, _I32, _VP, _VP]),
    ("rts_index_query_lines_d", _I32, [_VP, _I32, _VP, _U64, _I32, _VP, _VP]),
    ("rts_index_query_static", _I32, [_VP, _I32, _I32, _VP, _U64, _I32, _VP, _VP]),
    ("rts_index_size", _U64, [_VP]),
    ("rts_index_tree_view", _I32, [_VP, _VP, ctypes.POINTER(rts_tree_view)]),
    ("rts_index_query_plan", _I32, [_VP, _I32, _I32, _U64, ctypes.POINTER(rts_query_plan)]),
    ("rts_index_order_queries", _I32, [_VP, _VP, _U64, _I32, _VP, _VP]),
    ("rts_index_build", _I32, [_VP, _VP]),
    ("rts_index_optimize", _I32, [_VP, _VP]),
    ("rts_index_sync", _I32, [_VP, _VP]),
    ("rts_index_get_stats", _I32, [_VP, ctypes.POINTER(rts_index_stats)]),
    ("rts_index_memory", _I32, [_VP, ctypes.POINTER(rts_memory_usage)]),
    ("rts_sort_pairs", _I32, [_VP, _VP, _VP, _VP, _U64, _I32, _VP]),
    ("rts_tune_config_default", None, [ctypes.POINTER(rts_tune_config)]),
    ("rts_index_autotune", _I32, [_VP, _I32, _VP, _U64, _I32, _I32, ctypes.POINTER(rts_tune_config),
                                  ctypes.POINTER(rts_tune_result), _VP]),
    ("rts_serve_config_default", None, [ctypes.POINTER(rts_serve_config)]),
    ("rts_serve_run", _I32, [_VP, ctypes.POINTER(rts_serve_config), ctypes.POINTER(rts_request),
                             _U64, ctypes.POINTER(rts_update_batch), _U64,
                             ctypes.POINTER(rts_request_result), ctypes.POINTER(rts_serve_stats)]),
    ("rts_profile_enable", None, [_I32]),
    ("rts_profile_read", _I32, [ctypes.POINTER(rts_profile), _I32]),
    ("rts_last_error", ctypes.c_char_p, []),
    ("rts_version", ctypes.c_char_p, []),
]

_LIB = None


class RtsError(RuntimeError):
    pass


def lib():
    """Loads the HIP library (raises if it has not been built)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `python -m rtspatial_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def check(rc, what=""):
    if rc != RTS_OK:
        msg = lib().rts_last_error().decode(errors="replace")
        raise RtsError(f"{what} failed (status {rc}): {msg}")
