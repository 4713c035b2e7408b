"""Seeded synthetic workloads for the five benchmark configurations.

Each generator returns numpy float32 arrays: boxes [n, 4] as (x0, y0, x1, y1)
and points [n, 2]. Seeds: boxes 1000*c+1, queries 1000*c+2, mutations 1000*c+3,
request arrivals 1000*c+4 (c = configuration number), so that box and query
streams never coincide.

  c1  point query, small: 100k boxes (corner U[0,1)^2, sides U[0,0.01],
      clipped at 1 like test/test_commons.h:42-44) x 100k points U[0,1)^2
  c2  range-within: 1M boxes (sides U[0,1e-3]) x 1M query envelopes (sides U[0,0.01])
  c3  range-intersects, skewed: 10M boxes, centres from a 16-component Gaussian
      mixture (means U[0.1,0.9]^2, sigma 0.02, clipped to [0,1)), log-normal
      sides (mu ln 1e-4, sigma 1); 1M queries, same centres, sides LN(ln 5e-3, 1)
  c4  mutable index: c2-style 10M boxes; rounds of 1M deletes, 1M inserts,
      1M updates shifted by U[-1e-3,1e-3], then 1M point queries
  c5  served mixed stream over the c3 index: Poisson arrivals, batch sizes
      log-uniform in [1, 1e6], 40% point / 30% within / 30% intersects, and
      100k-box update batches at 10 Hz
"""
import math

import numpy as np

ONE_BELOW = np.nextafter(np.float32(1.0), np.float32(0.0))


def _rng(cfg, stream):
    return np.random.default_rng(1000 * cfg + stream)


def uniform_boxes(rng, n, max_side):
    """Corner U[0,1)^2, sides U[0, max_side], max corner clipped at 1 (FP32)."""
    x0 = rng.random(n, dtype=np.float32)
    y0 = rng.random(n, dtype=np.float32)
    w = (rng.random(n, dtype=np.float32) * np.float32(max_side)).astype(np.float32)
    h = (rng.random(n, dtype=np.float32) * np.float32(max_side)).astype(np.float32)
    x1 = np.minimum(x0 + w, np.float32(1.0))
    y1 = np.minimum(y0 + h, np.float32(1.0))
    return np.stack([x0, y0, x1, y1], axis=1).astype(np.float32)


def uniform_points(rng, n):
    return rng.random((n, 2), dtype=np.float32)


def mixture_means(rng, k=16):
    return rng.uniform(0.1, 0.9, size=(k, 2))


def mixture_boxes(rng, means, n, side_mu, side_sigma=1.0, sigma=0.02, chunk=1 << 22,
                  dtype=np.float32):
    """Centres from the Gaussian mixture, clipped to [0, 1), log-normal sides;
    box = centre -+ side/2 in `dtype` (float64: the same draws, kept in double,
    so the coordinates are not floats)."""
    out = np.empty((n, 4), dtype)
    for s in range(0, n, chunk):
        m = min(chunk, n - s)
        comp = rng.integers(0, means.shape[0], size=m)
        c = means[comp] + rng.normal(0.0, sigma, size=(m, 2))
        c = np.clip(c, 0.0, float(ONE_BELOW)).astype(dtype)
        side = rng.lognormal(side_mu, side_sigma, size=(m, 2)).astype(dtype)
        half = (side * dtype(0.5)).astype(dtype)
        out[s:s + m, 0] = c[:, 0] - half[:, 0]
        out[s:s + m, 1] = c[:, 1] - half[:, 1]
        out[s:s + m, 2] = c[:, 0] + half[:, 0]
        out[s:s + m, 3] = c[:, 1] + half[:, 1]
    return out


def config1(n_boxes=100_000, n_points=100_000):
    b = uniform_boxes(_rng(1, 1), n_boxes, 0.01)
    q = uniform_points(_rng(1, 2), n_points)
    return b, q


def config2(n_boxes=1_000_000, n_queries=1_000_000):
    b = uniform_boxes(_rng(2, 1), n_boxes, 1e-3)
    q = uniform_boxes(_rng(2, 2), n_queries, 0.01)
    return b, q


def config3(n_boxes=10_000_000, n_queries=1_000_000, dtype=np.float32):
    rb = _rng(3, 1)
    means = mixture_means(rb)
    b = mixture_boxes(rb, means, n_boxes, math.log(1e-4), dtype=dtype)
    q = mixture_boxes(_rng(3, 2), means, n_queries, math.log(5e-3), dtype=dtype)
    return b, q


class MutationRounds:
    """c4: the initial boxes and, per round, (delete ids, inserted boxes,
    update ids, updated boxes, query points). Ids are drawn among the boxes
    alive at the start of the round; updates avoid the round's deletes."""

    def __init__(self, n_boxes=10_000_000, n_ops=1_000_000, n_queries=1_000_000):
        self.n_ops = n_ops
        self.n_queries = n_queries
        self.initial = uniform_boxes(_rng(4, 1), n_boxes, 1e-3)
        self.rng = _rng(4, 3)
        self.qrng = _rng(4, 2)
        self.alive = np.ones(n_boxes, dtype=bool)
        self.boxes = self.initial.copy()

    def next_round(self):
        rng = self.rng
        live = np.flatnonzero(self.alive)
        pick = rng.choice(live.size, size=2 * self.n_ops, replace=False)
        del_ids = live[pick[: self.n_ops]].astype(np.int64)
        upd_ids = live[pick[self.n_ops:]].astype(np.int64)
        ins = uniform_boxes(rng, self.n_ops, 1e-3)
        shift = rng.uniform(-1e-3, 1e-3, size=(self.n_ops, 2)).astype(np.float32)
        upd = self.boxes[upd_ids].copy()
        upd[:, 0] += shift[:, 0]
        upd[:, 2] += shift[:, 0]
        upd[:, 1] += shift[:, 1]
        upd[:, 3] += shift[:, 1]
        pts = uniform_points(self.qrng, self.n_queries)
        # Apply to the host mirror in the index's order: delete, insert, update.
        self.alive[del_ids] = False
        self.boxes = np.concatenate([self.boxes, ins])
        self.alive = np.concatenate([self.alive, np.ones(self.n_ops, dtype=bool)])
        self.boxes[upd_ids] = upd
        return del_ids, ins, upd_ids, upd, pts


def request_trace(n_requests, rate_hz, seed_cfg=5, max_batch=1_000_000,
                  mix=(0.4, 0.3, 0.3)):
    """c5 arrivals: (arrival time s, kind, size) with kind 0 = point,
    1 = within, 2 = intersects; sizes round(exp(U[0, ln max_batch]))."""
    rng = _rng(seed_cfg, 4)
    gaps = rng.exponential(1.0 / rate_hz, size=n_requests)
    t = np.cumsum(gaps)
    kind = rng.choice(3, size=n_requests, p=list(mix))
    size = np.rint(np.exp(rng.uniform(0.0, math.log(max_batch), size=n_requests))).astype(np.int64)
    size = np.maximum(size, 1)
    return t, kind, size


class ServedStream:
    """c5: the served mixed stream over the 10M-box Gaussian-mixture index.

    Boxes: n_boxes mixture boxes (seed 3001, sides LN(ln 1e-4, 1)). Query
    pools: pool_n envelopes and pool_n points from the same mixture (seed
    5002, sides LN(ln 5e-3, 1)). Requests (seed 5004): kind 0 = point
    (contains), 1 = within, 2 = intersects with probabilities `mix`, size
    round(exp(U[0, ln max_request])), each a contiguous slice of its pool at a
    random offset, arrival gaps exponential with mean 1 (divided by the
    offered rate). Update batches (seed 5003): batch b gives the next
    `update_size` ids of one random permutation of the boxes their original
    box shifted by U[-1e-3, 1e-3] per axis; they arrive at 10 Hz."""

    def __init__(self, n_boxes=10_000_000, n_requests=600, max_request=1_000_000,
                 pool_n=1 << 20, mix=(0.4, 0.3, 0.3), n_update_batches=100,
                 update_size=100_000):
        rb = np.random.default_rng(3001)
        self.means = mixture_means(rb)
        self.boxes = mixture_boxes(rb, self.means, n_boxes, math.log(1e-4))
        rq = np.random.default_rng(5002)
        self.pool_n = pool_n
        self.env_pool = mixture_boxes(rq, self.means, pool_n, math.log(5e-3))
        self.pt_pool = mixture_boxes(rq, self.means, pool_n, math.log(5e-3))[:, :2].copy()
        rt = np.random.default_rng(5004)
        self.kinds = rt.choice(3, size=n_requests, p=list(mix))
        self.sizes = np.maximum(
            1, np.rint(np.exp(rt.uniform(0, math.log(max_request), n_requests)))).astype(np.int64)
        self.sizes = np.minimum(self.sizes, pool_n)
        self.offsets = rt.integers(0, pool_n - self.sizes + 1)
        self.gaps = rt.exponential(1.0, size=n_requests)
        ru = np.random.default_rng(5003)
        perm = ru.permutation(n_boxes)
        self.updates = []
        for b in range(n_update_batches):
            s = (b * update_size) % n_boxes
            ids = np.sort(perm[s:s + update_size]).astype(np.int64)
            sh = ru.uniform(-1e-3, 1e-3, size=(ids.shape[0], 2)).astype(np.float32)
            nb = self.boxes[ids].copy()
            nb[:, [0, 2]] += sh[:, :1]
            nb[:, [1, 3]] += sh[:, 1:]
            self.updates.append((ids, nb))

    def predicate(self, i):
        """Oracle predicate code of request i: contains (points), within, intersects."""
        return (1, 2, 0)[int(self.kinds[i])]

    def queries(self, i):
        o, n = int(self.offsets[i]), int(self.sizes[i])
        return self.pt_pool[o:o + n] if self.kinds[i] == 0 else self.env_pool[o:o + n]

    def arrivals(self, rate):
        """Arrival times at `rate` requests per second (0: all at t = 0)."""
        return np.cumsum(self.gaps) / rate if rate else np.zeros(self.gaps.shape[0])

    def update_times(self, horizon_s):
        """Arrival times of the update batches that arrive within horizon_s."""
        n = min(len(self.updates), int(horizon_s * 10) + 1)
        return [0.1 * i for i in range(n)]
