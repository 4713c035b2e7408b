"""Benchmark of the LBVH spatial index on one MI355X.

Default workload (the 10M-box index the headline metric is quoted on):
range-intersects over 10M Gaussian-mixture boxes with 1M skewed query
envelopes; one step = one Query(kIntersects) over the 1M queries with the
collecting handler, every result pair written to a resident device queue.
Prints one JSON line.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Mqueries/s + p50/p99 latency, 10M-box index: served/batched vs plain LBVH"
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_threads():
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        return int(env)
    return max(1, min(16, os.cpu_count() or 1))


def query_bytes(nq, s_q, pairs, n_boxes):
    """Algorithmic HBM bytes of one query launch: read the queries, write the
    8-byte pairs, and one root-to-leaf descent per query capped at reading the
    whole tree once (48 bytes per box)."""
    descent = nq * math.ceil(math.log2(max(n_boxes, 2))) * 32
    return nq * s_q + pairs * 8 + min(descent, 48 * n_boxes)


def cpu_baseline(boxes, queries, pred, threads, sample_rtree=20000, sample_bf=64):
    """The oracle's STR R-tree (bulk load + parallel queries) and brute-force
    scan, timed on the host on a bounded sample of the same workload."""
    from oracle import oracle as O

    rng = np.random.default_rng(99)
    t0 = time.perf_counter()
    tree = O.RTree(boxes)
    t_build = time.perf_counter() - t0
    qs = queries[rng.choice(queries.shape[0], min(sample_rtree, queries.shape[0]), replace=False)]
    tree.count(pred, qs[:256], threads)  # warm-up
    t0 = time.perf_counter()
    n_rt = tree.count(pred, qs, threads)
    t_rt = time.perf_counter() - t0
    qb = qs[:sample_bf]
    t0 = time.perf_counter()
    n_bf = O.bf_count(pred, boxes, qb, threads=threads)
    t_bf = time.perf_counter() - t0
    del tree
    return {
        "value": round(qs.shape[0] / t_rt / 1e6, 6),
        "unit": "Mqueries/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"oracle STR R-tree (fan-out 16) over all {boxes.shape[0]} boxes, "
                   f"{qs.shape[0]} random queries of the workload ({n_rt} pairs, counted); "
                   f"R-tree bulk load {boxes.shape[0] / t_build / 1e6:.3f} Mboxes/s single-threaded; "
                   f"brute-force scan of {qb.shape[0]} queries x {boxes.shape[0]} boxes: "
                   f"{qb.shape[0] / t_bf / 1e6:.6f} Mqueries/s ({n_bf} pairs)"),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--boxes", type=int, default=10_000_000)
    ap.add_argument("--queries", type=int, default=1_000_000)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; run with --gpus 1",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import torch

    import rtspatial_amd as R
    from rtspatial_amd import workloads as W

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)

    log(f"generating workload: {args.boxes} boxes, {args.queries} queries")
    boxes_np, queries_np = W.config3(args.boxes, args.queries)
    boxes = torch.from_numpy(boxes_np).to(dev)
    queries = torch.from_numpy(queries_np).to(dev)
    pred = R.Predicate.kIntersects

    ix = R.SpatialIndex()
    ix.Init(R.Config(max_geometries=args.boxes))

    # Build rate: rebuild from the same boxes a few times, phase-timed.
    R.profile_enable(True)
    build_ms = []
    for rep in range(3):
        ix.Clear()
        ix.Insert(boxes)
        torch.cuda.synchronize()
        R.profile_read(reset=True)
        ix.Build()
        ix.Sync()
        p = R.profile_read(reset=True)
        build_ms.append(sum(p[k][1] for k in ("build_morton", "build_sort", "build_tree")))
        sort_ms = p["build_sort"][1]
    build_ms_med = float(np.median(build_ms))
    log(f"build: {build_ms_med:.3f} ms ({args.boxes / build_ms_med / 1e3:.1f} Mboxes/s), sort {sort_ms:.3f} ms")

    cnt = R.SharedValue()
    cnt.set(0)
    ix.Query(pred, queries, cnt)
    n_pairs = cnt.get()
    log(f"pairs per step: {n_pairs}")
    queue = R.Queue(n_pairs + 1024)

    for _ in range(args.warmup):
        queue.Clear()
        ix.Query(pred, queries, queue)
    torch.cuda.synchronize()
    assert queue.size() == n_pairs, (queue.size(), n_pairs)

    R.profile_read(reset=True)
    starts = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        starts[k].record()
        queue.Clear()
        ix.Query(pred, queries, queue)
        ends[k].record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    prof = R.profile_read(reset=True)
    R.profile_enable(False)
    ix.Sync()
    assert queue.size() == n_pairs and not queue.overflowed()

    step_ms = np.array([s.elapsed_time(e) for s, e in zip(starts, ends)])
    ms_per_step = wall * 1e3 / args.steps
    nq = queries.shape[0]
    value = nq / (ms_per_step / 1e3) / 1e6
    trav_n, trav_ms = prof["traverse"]
    trav_avg = trav_ms / max(trav_n, 1)
    order_ms = prof["query_order"][1] / max(prof["query_order"][0], 1)
    qbytes = query_bytes(nq, 16, n_pairs, args.boxes)
    achieved = qbytes / (trav_avg / 1e3) / 1e9

    out = {
        "metric": METRIC,
        "value": round(value, 4),
        "unit": "Mqueries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": (f"range-intersects, skewed: {args.boxes} Gaussian-mixture boxes (16 "
                         f"components, sigma 0.02, log-normal sides mu=ln 1e-4) x {nq} query "
                         f"envelopes (log-normal sides mu=ln 5e-3); plain per-batch LBVH, one "
                         f"Query(kIntersects) per step, all pairs collected on the device"),
            "path": "plain",
            "boxes": args.boxes,
            "queries": nq,
            "pairs_per_step": n_pairs,
            "seeds": {"boxes": 3001, "queries": 3002},
        },
        "latency_ms": {"p50": round(float(np.percentile(step_ms, 50)), 4),
                       "p99": round(float(np.percentile(step_ms, 99)), 4)},
        "build": {"ms": round(build_ms_med, 4), "mboxes_per_s": round(args.boxes / build_ms_med / 1e3, 3),
                  "sort_ms": round(sort_ms, 4), "sort_mkeys_per_s": round(args.boxes / sort_ms / 1e3, 3)},
        "phases_ms_per_step": {"query_order": round(order_ms, 4), "traverse": round(trav_avg, 4)},
        "roofline": {
            "bound": "hbm",
            "kernel": "k_query_thread<intersects, envelopes, collect>",
            "achieved": round(achieved, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": None,
            "algorithmic_bytes_per_launch": int(qbytes),
        },
    }
    if not args.no_cpu_baseline:
        log("cpu baseline (oracle R-tree / brute force on a sample)")
        out["cpu_baseline"] = cpu_baseline(boxes_np, queries_np, int(pred), cpu_threads())
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
