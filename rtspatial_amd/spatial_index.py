"""Python mirror of rtspatial::SpatialIndex<float, 2> and <double, 2> over the C ABI.

Reference: include/rtspatial/spatial_index.cuh:20-931 (Config, SpatialIndex),
utils/queue.h:60-113 (Queue), utils/shared_value.h:11-58 (SharedValue),
geom/predicates.h:7-10 (Predicate). Method names and argument meanings follow
the reference; tensors are torch tensors on the GPU (or pinned host tensors,
as the reference's ArrayView of a pinned_vector).

  points    float32 [n, 2]        Point<float,2>
  envelopes float32 [n, 4]        Envelope: (min x, min y, max x, max y)
  ids       int64   [n]           size_t ids (Delete)
  updates   (ids int64 [n], envelopes float32 [n, 4])   thrust::pair<size_t, envelope_t>
  results   int32   [k, 2]        (geom_id, query_id), collecting_handlers.h:9-31

SpatialIndex(torch.float64) is SpatialIndex<double, 2>: points, envelopes and
updates carry float64 coordinates (updates: int64 [n, 5], see pack_updates).
Its predicates are exact in double precision; Config.coord =
RTS_COORD_F64_AS_F32 instead rounds coordinates to the nearest float, which
reproduces the reference's results (its Point<double,2>::get_coordinate
returns float, geom/point.cuh:68).
"""
import ctypes
import dataclasses
import enum

import torch

from . import _native as N


class Predicate(enum.IntEnum):
    """geom/predicates.h:7-10, plus kWithin (data box inside the query)."""
    kIntersects = N.RTS_INTERSECTS
    kContains = N.RTS_CONTAINS
    kWithin = N.RTS_WITHIN


@dataclasses.dataclass
class Config:
    """Config (spatial_index.cuh:22-37). ptx_root, prefer_fast_build_* and
    compact configure OptiX in the reference and are accepted and ignored."""
    ptx_root: str = ""
    max_geometries: int = 10_000_000
    max_queries: int = 1_000_000
    preallocate: bool = False
    prefer_fast_build_geom: bool = False
    prefer_fast_build_query: bool = False
    compact: bool = False
    geom_sample_rate: float = 0.2
    query_sample_rate: float = 0.2
    intersect_cost_weight: float = 0.95
    max_geom_samples: int = 10000
    max_query_samples: int = 1000
    max_parallelism: int = 512
    # MI355X additions
    leaf_size: int = 8
    rebuild_policy: int = N.RTS_REBUILD_SAH
    sah_rebuild_ratio: float = 1.5
    traversal: int = N.RTS_TRAVERSAL_AUTO
    sort_queries: int = 0
    # Double indexes: RTS_COORD_F64 (exact, the default) or RTS_COORD_F64_AS_F32.
    # The index's dtype decides between float and double.
    coord: int = N.RTS_COORD_F64
    # Inserts go to a small second tree while they are at most this percentage
    # of the main tree (0: every Insert rebuilds the whole tree).
    insert_delta_pct: int = 2

    def to_c(self):
        c = N.rts_config()
        N.lib().rts_config_default(ctypes.byref(c))
        for f, _ in N.rts_config._fields_:
            v = getattr(self, f)
            setattr(c, f, int(v) if isinstance(v, bool) else v)
        return c


def _stream_handle(stream):
    if stream is None:
        stream = torch.cuda.current_stream()
    if isinstance(stream, torch.cuda.Stream):
        return ctypes.c_void_p(stream.cuda_stream)
    return ctypes.c_void_p(int(stream))


def _handler_arg(arg, what):
    if isinstance(arg, SharedValue):
        return N.RTS_HANDLER_COUNT, arg.data()
    if isinstance(arg, Queue):
        return N.RTS_HANDLER_COLLECT, arg.DeviceObject()
    raise TypeError(f"{what}: arg must be a SharedValue or a Queue")


def _check_array(t, cols, dtype, what):
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{what}: expected a torch tensor")
    if t.dtype != dtype:
        raise TypeError(f"{what}: expected {dtype}, got {t.dtype}")
    if cols is not None and (t.dim() != 2 or t.shape[1] != cols):
        raise ValueError(f"{what}: expected shape [n, {cols}], got {tuple(t.shape)}")
    if not t.is_contiguous():
        raise ValueError(f"{what}: tensor must be contiguous")
    if not (t.is_cuda or t.is_pinned()):
        raise ValueError(f"{what}: tensor must be on the GPU or in pinned host memory")
    return t


class SharedValue:
    """A device-resident unsigned 64-bit counter (utils/shared_value.h:11-58),
    the `arg` of the counting handler (counting_handlers.h:7-25)."""

    def __init__(self, device=None):
        self._t = torch.zeros(1, dtype=torch.int64, device=device or "cuda")

    def data(self):
        return self._t.data_ptr()

    def set(self, value, stream=None):
        with torch.cuda.stream(stream) if isinstance(stream, torch.cuda.Stream) else _null():
            self._t.fill_(int(value))

    def get(self, stream=None):
        if isinstance(stream, torch.cuda.Stream):
            stream.synchronize()
        else:
            torch.cuda.synchronize()
        return int(self._t.item()) & 0xFFFFFFFFFFFFFFFF

    @property
    def tensor(self):
        return self._t


class Queue:
    """Result queue of (geom_id, query_id) pairs (utils/queue.h:60-113) with
    a 64-bit counter. DeviceObject() is the device pointer the collecting
    handler takes as `arg`."""

    def __init__(self, capacity=0, device=None):
        self._device = torch.device(device or "cuda")
        self._counter = torch.zeros(1, dtype=torch.int64, device=self._device)
        self._desc = None
        self._data = None
        self.Init(capacity)

    def Init(self, capacity):
        capacity = max(int(capacity), 1)
        self._data = torch.empty((capacity, 2), dtype=torch.int32, device=self._device)
        host = N.rts_dev_queue(self._data.data_ptr(), capacity, self._counter.data_ptr())
        raw = torch.frombuffer(bytearray(bytes(host)), dtype=torch.uint8)
        self._desc = raw.to(self._device)
        self._counter.zero_()

    def Clear(self, stream=None):
        with torch.cuda.stream(stream) if isinstance(stream, torch.cuda.Stream) else _null():
            self._counter.zero_()

    def size(self, stream=None):
        """Number of results reported (may exceed capacity on overflow)."""
        if isinstance(stream, torch.cuda.Stream):
            stream.synchronize()
        else:
            torch.cuda.synchronize()
        return int(self._counter.item())

    def capacity(self):
        return self._data.shape[0]

    def overflowed(self, stream=None):
        return self.size(stream) > self.capacity()

    def DeviceObject(self):
        return self._desc.data_ptr()

    def pairs(self, stream=None):
        """The stored (geom_id, query_id) pairs as an int32 [k, 2] tensor."""
        k = min(self.size(stream), self.capacity())
        return self._data[:k]

    @property
    def counter(self):
        return self._counter


class _null:
    def __enter__(self):
        return self

    def __exit__(self, *a):
        return False


class SpatialIndex:
    """rtspatial::SpatialIndex<COORD_T, 2> (spatial_index.cuh:62-931);
    COORD_T is torch.float32 (default) or torch.float64."""

    def __init__(self, coord_t=torch.float32):
        if coord_t not in (torch.float32, torch.float64):
            raise TypeError("SpatialIndex: coord_t must be torch.float32 or torch.float64")
        self._h = None
        self.coord_t = coord_t
        self._d = coord_t == torch.float64

    # -- lifecycle ---------------------------------------------------------
    def Init(self, config=None):
        """spatial_index.cuh:83-110."""
        self._destroy()
        cfg = (config or Config()).to_c()
        if not self._d:
            cfg.coord = N.RTS_COORD_F32
        elif cfg.coord not in (N.RTS_COORD_F64, N.RTS_COORD_F64_AS_F32):
            raise ValueError("Init: a double index takes coord RTS_COORD_F64 or RTS_COORD_F64_AS_F32")
        h = ctypes.c_void_p()
        N.check(N.lib().rts_index_create(ctypes.byref(cfg), ctypes.byref(h)), "Init")
        self._h = h
        self.config = config or Config()

    def Clear(self):
        """spatial_index.cuh:118-129."""
        N.check(N.lib().rts_index_clear(self._handle()), "Clear")

    def _destroy(self):
        if self._h is not None:
            N.lib().rts_index_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self._destroy()
        except Exception:
            pass

    def _handle(self):
        if self._h is None:
            raise RuntimeError("SpatialIndex.Init() has not been called")
        return self._h

    # -- mutation ------------------------------------------------------------
    def Insert(self, envelopes, stream=None):
        """spatial_index.cuh:142-196: appends boxes with ids size()..size()+n-1."""
        _check_array(envelopes, 4, self.coord_t, "Insert")
        fn = N.lib().rts_index_insert_d if self._d else N.lib().rts_index_insert
        N.check(fn(self._handle(), envelopes.data_ptr(), envelopes.shape[0],
                   _stream_handle(stream)), "Insert")

    def Delete(self, ids, stream=None):
        """spatial_index.cuh:208-265. ids >= size() are ignored."""
        _check_array(ids, None, torch.int64, "Delete")
        N.check(N.lib().rts_index_delete(self._handle(), ids.data_ptr(), ids.numel(),
                                         _stream_handle(stream)), "Delete")

    def Update(self, updates, stream=None):
        """spatial_index.cuh:280-336. `updates` is a packed int64 [n, 3] tensor
        ([n, 5] for a double index) from pack_updates(), or an (ids,
        envelopes) tuple."""
        if isinstance(updates, tuple):
            updates = pack_updates(*updates, coord_t=self.coord_t)
        _check_array(updates, 5 if self._d else 3, torch.int64, "Update")
        fn = N.lib().rts_index_update_d if self._d else N.lib().rts_index_update
        N.check(fn(self._handle(), updates.data_ptr(), updates.shape[0], _stream_handle(stream)),
                "Update")

    # -- queries -------------------------------------------------------------
    def Query(self, predicate, queries, arg, stream=None):
        """spatial_index.cuh:352-391 (points, [n, 2]) and :407-515 (envelopes,
        [n, 4]). `arg` is a SharedValue (counting handler) or a Queue
        (collecting handler)."""
        _check_array(queries, None, self.coord_t, "Query")
        if queries.dim() != 2 or queries.shape[1] not in (2, 4):
            raise ValueError("Query: queries must be [n, 2] points or [n, 4] envelopes")
        handler, ptr = _handler_arg(arg, "Query")
        L = N.lib()
        if self._d:
            fn = L.rts_index_query_points_d if queries.shape[1] == 2 else L.rts_index_query_envelopes_d
        else:
            fn = L.rts_index_query_points if queries.shape[1] == 2 else L.rts_index_query_envelopes
        N.check(fn(self._handle(), int(predicate), queries.data_ptr(), queries.shape[0], handler,
                   ptr, _stream_handle(stream)), "Query")

    def QueryLines(self, predicate, segments, arg, stream=None):
        """spatial_index.cuh:517-547: Query(p, ArrayView<line_t>, arg) for
        segments [n, 4] = (x1, y1, x2, y2). INTERSECTS reports every data box
        the closed segment meets (exact); CONTAINS reports nothing, as the
        reference's; WITHIN is an error."""
        _check_array(segments, 4, self.coord_t, "QueryLines")
        handler, ptr = _handler_arg(arg, "QueryLines")
        L = N.lib()
        fn = L.rts_index_query_lines_d if self._d else L.rts_index_query_lines
        N.check(fn(self._handle(), int(predicate), segments.data_ptr(), segments.shape[0], handler,
                   ptr, _stream_handle(stream)), "QueryLines")

    # -- extras ----------------------------------------------------------------
    def size(self):
        return int(N.lib().rts_index_size(self._handle()))

    def Build(self, stream=None):
        N.check(N.lib().rts_index_build(self._handle(), _stream_handle(stream)), "Build")

    def Sync(self, stream=None):
        N.check(N.lib().rts_index_sync(self._handle(), _stream_handle(stream)), "Sync")

    def Autotune(self, predicate, queries, collect=True, sample=32768, leaf_sizes=(4, 8, 16, 32),
                 reps=3, apply=True, stream=None):
        """Times the query over a strided sample of `queries` under every
        plan (leaf size, per-lane or per-wavefront traversal, Morton or input
        order) and, with `apply`, keeps the fastest for later queries of this
        predicate and query kind. Returns the choice and the sample timings."""
        _check_array(queries, None, self.coord_t, "Autotune")
        if queries.dim() != 2 or queries.shape[1] not in (2, 4):
            raise ValueError("Autotune: queries must be [n, 2] points or [n, 4] envelopes")
        c = N.rts_tune_config()
        N.lib().rts_tune_config_default(ctypes.byref(c))
        c.sample = int(sample)
        c.leaf_sizes = sum(1 << (int(ls).bit_length() - 1) for ls in leaf_sizes)
        c.reps = int(reps)
        c.apply = 1 if apply else 0
        r = N.rts_tune_result()
        handler = N.RTS_HANDLER_COLLECT if collect else N.RTS_HANDLER_COUNT
        N.check(N.lib().rts_index_autotune(self._handle(), int(predicate), queries.data_ptr(),
                                           queries.shape[0], 1 if queries.shape[1] == 2 else 0,
                                           handler, ctypes.byref(c), ctypes.byref(r),
                                           _stream_handle(stream)), "Autotune")
        return {f: getattr(r, f) for f, _ in N.rts_tune_result._fields_}

    def Optimize(self, stream=None):
        """spatial_index.cuh:766-826 merges the per-Insert BVHs; here it brings
        the tree up to date and merges the small tree of recent inserts into
        the main one (on `stream`)."""
        N.check(N.lib().rts_index_optimize(self._handle(), _stream_handle(stream)), "Optimize")

    def memory(self):
        """Device bytes held: geometries, aabbs (leaf boxes), bvh, total."""
        m = N.rts_memory_usage()
        N.check(N.lib().rts_index_memory(self._handle(), ctypes.byref(m)), "memory")
        return {f: getattr(m, f) for f, _ in N.rts_memory_usage._fields_}

    def PrintMemoryUsage(self):
        """spatial_index.cuh:893-903: one line, in MB."""
        m = self.memory()
        mb = {k: v // (1024 * 1024) for k, v in m.items()}
        print(f"BVH {mb['bvh']} MB, Geometries {mb['geometries']} MB, AABBs {mb['aabbs']} MB, "
              f"Total {mb['total']} MB", flush=True)

    def stats(self):
        s = N.rts_index_stats()
        N.check(N.lib().rts_index_get_stats(self._handle(), ctypes.byref(s)), "stats")
        return {f: getattr(s, f) for f, _ in N.rts_index_stats._fields_}


def pack_updates(ids, envelopes, coord_t=torch.float32):
    """(ids int64 [n], envelopes [n, 4]) -> int64 [n, 3], the 24-byte
    rts_update records (thrust::pair<size_t, envelope_t>), or, with coord_t
    float64, int64 [n, 5], the 40-byte rts_update_d records."""
    ids = ids.reshape(-1, 1).to(torch.int64)
    words = 2 if coord_t == torch.float32 else 4
    env = envelopes.to(coord_t).contiguous().view(torch.int64).reshape(-1, words)
    return torch.cat([ids, env], dim=1).contiguous()


def sort_pairs(keys, vals=None, key_bits=32, stream=None):
    """Stable onesweep radix sort of uint32 keys (int32 tensors) with uint32
    values (identity when vals is None). Returns (sorted_keys, sorted_vals)."""
    n = keys.numel()
    ko = torch.empty_like(keys)
    vo = torch.empty(n, dtype=torch.int32, device=keys.device)
    N.check(N.lib().rts_sort_pairs(keys.data_ptr(), vals.data_ptr() if vals is not None else None,
                                   ko.data_ptr(), vo.data_ptr(), n, key_bits,
                                   _stream_handle(stream)), "sort_pairs")
    return ko, vo


def profile_enable(on=True):
    """Per-phase GPU timing (HIP events on the caller's stream)."""
    N.lib().rts_profile_enable(1 if on else 0)


def profile_read(reset=True):
    """{phase: (launch count, total ms)} since the last reset."""
    p = N.rts_profile()
    N.check(N.lib().rts_profile_read(ctypes.byref(p), 1 if reset else 0), "profile_read")
    return {name: (int(p.count[i]), float(p.ms[i])) for i, name in enumerate(N.RTS_PHASES)}
