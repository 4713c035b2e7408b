"""Served / batched execution of a request trace (rts_serve_run).

A request is (arrival time, predicate, queries tensor [n, 2] points or [n, 4]
envelopes -- or segments with `segments=True` -- float32, or float64 on an
exact double index, optional output capacity). `run()` returns one dict per request
(admission and completion times, result count, epoch, re-launches, and the
pairs as an int32 [k, 2] (geom_id, query index) tensor when collected) plus
run statistics.
"""
import ctypes
import dataclasses

import torch

from . import _native as N
from .spatial_index import Predicate, pack_updates


@dataclasses.dataclass
class Request:
    queries: torch.Tensor
    predicate: int = Predicate.kIntersects
    arrival_s: float = 0.0
    collect: int = 0  # capacity of the output buffer (0: results are dropped)
    segments: bool = False  # [n, 4] rows are line segments (x1, y1, x2, y2), not envelopes


@dataclasses.dataclass
class UpdateBatch:
    ids: torch.Tensor
    boxes: torch.Tensor
    arrival_s: float = 0.0


def config(mode="batched", **kw):
    c = N.rts_serve_config()
    N.lib().rts_serve_config_default(ctypes.byref(c))
    c.mode = N.RTS_SERVE_PLAIN if mode == "plain" else N.RTS_SERVE_BATCHED
    for k, v in kw.items():
        if not hasattr(c, k):
            raise AttributeError(f"rts_serve_config has no field {k}")
        setattr(c, k, int(v))
    return c


def run(index, requests, updates=(), mode="batched", **cfg):
    """Runs the trace; requests/updates must be sorted by arrival time."""
    c = config(mode, **cfg)
    n = len(requests)
    reqs = (N.rts_request * max(n, 1))()
    outs = []
    for i, r in enumerate(requests):
        q = r.queries
        want = getattr(index, "coord_t", torch.float32)
        if (q.dtype != want or q.dim() != 2 or q.shape[1] not in (2, 4) or not q.is_contiguous() or
                (r.segments and q.shape[1] != 4)):
            raise ValueError(f"request queries must be contiguous {want} [n, 2] or [n, 4] "
                             "(the index's coordinate type)")
        out = torch.empty((max(r.collect, 1), 2), dtype=torch.int32, device=q.device) if r.collect else None
        outs.append(out)
        reqs[i].arrival_s = float(r.arrival_s)
        reqs[i].queries = q.data_ptr() if q.numel() else None
        reqs[i].n = q.shape[0]
        reqs[i].points = 2 if r.segments else (1 if q.shape[1] == 2 else 0)
        reqs[i].predicate = int(r.predicate)
        reqs[i].out = out.data_ptr() if out is not None else None
        reqs[i].out_capacity = r.collect
    nu = len(updates)
    ups = (N.rts_update_batch * max(nu, 1))()
    keep = []
    for i, u in enumerate(updates):
        packed = pack_updates(u.ids, u.boxes, getattr(index, "coord_t", torch.float32))
        keep.append(packed)
        ups[i].arrival_s = float(u.arrival_s)
        ups[i].updates = packed.data_ptr()
        ups[i].n = packed.shape[0]
    torch.cuda.synchronize()
    res = (N.rts_request_result * max(n, 1))()
    st = N.rts_serve_stats()
    N.check(N.lib().rts_serve_run(index._handle(), ctypes.byref(c), reqs, n, ups, nu, res,
                                  ctypes.byref(st)), "serve_run")
    torch.cuda.synchronize()
    results = []
    for i in range(n):
        r = res[i]
        d = {f: getattr(r, f) for f, _ in N.rts_request_result._fields_ if f != "pad"}
        if outs[i] is not None:
            d["pairs"] = outs[i][: min(r.count, requests[i].collect)]
        results.append(d)
    stats = {f: getattr(st, f) for f, _ in N.rts_serve_stats._fields_}
    return results, stats
