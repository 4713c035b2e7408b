"""Benchmarks of the LBVH spatial index on one MI355X. Prints one JSON line.

Default workload (--workload served): the served mixed stream over the 10M-box
index the headline metric is quoted on. 10M Gaussian-mixture boxes are held
resident; requests arrive as a Poisson stream with sizes log-uniform in
[1, 1e6] queries, 40% point / 30% within / 30% intersects, and 100k-box update
batches arrive at 10 Hz. The same trace is answered by the served path
(continuous batching: persistent work-stealing batches, up to 3 in flight)
and by the plain per-request path (one query call per request, in arrival
order). `value` is the served throughput (Mqueries/s) at the highest offered
load whose p99 request latency does not exceed the plain path's p99 at 70% of
the plain path's capacity; both capacities and latencies are reported.

--workload c3: one plain Query(kIntersects) of 1M skewed query envelopes over
the 10M boxes per step, all ~1e10 pairs collected on the device.
--workload c3f64: the same on a double-coordinate index (SpatialIndex<double, 2>,
exact double predicates), boxes and queries drawn in double.
--workload lines: line-segment intersects, the diagonals of the c3 query
envelopes as segments over the same boxes.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload served|c3|c3f64|lines|c4]
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


def _cgroup_cpus():
    """CPUs granted by the cgroup CPU quota (cpu.max), or None if unlimited."""
    for path in ("/sys/fs/cgroup/cpu.max",):
        try:
            with open(path) as f:
                quota, period = f.read().split()[:2]
            if quota != "max":
                return max(1, int(int(quota) // int(period)))
        except (OSError, ValueError):
            pass
    try:
        with open("/sys/fs/cgroup/cpu/cpu.cfs_quota_us") as f:
            q = int(f.read())
        with open("/sys/fs/cgroup/cpu/cpu.cfs_period_us") as f:
            per = int(f.read())
        if q > 0:
            return max(1, q // per)
    except (OSError, ValueError):
        pass
    return None


def cpu_info():
    """The host CPUs the CPU baselines run on: the model, the host's logical
    CPUs (os.cpu_count()), the CPUs this process may run on (affinity, cgroup
    quota) and the threads used: every CPU the process is granted, capped by
    OMP_NUM_THREADS when the job's environment sets it (the GPU machines set
    it to the job's CPU share)."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        affinity = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        affinity = os.cpu_count() or 1
    quota = _cgroup_cpus()
    threads = affinity if quota is None else min(affinity, quota)
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        threads = min(threads, int(env))
    return {"model": model, "host_logical_cpus": os.cpu_count(), "affinity_cpus": affinity,
            "cgroup_cpus": quota, "omp_num_threads": env, "threads": max(1, threads)}


def cpu_threads():
    return cpu_info()["threads"]


def _cores_note(ci):
    return (f"{ci['threads']} threads on {ci['model']} (host: {ci['host_logical_cpus']} logical "
            f"CPUs; this job may use {ci['affinity_cpus']} by affinity"
            f"{'' if ci['cgroup_cpus'] is None else ', ' + str(ci['cgroup_cpus']) + ' by cgroup quota'}"
            f"{'' if not ci['omp_num_threads'] else ', OMP_NUM_THREADS=' + ci['omp_num_threads']})")


def query_bytes(nq, s_q, pairs, n_boxes, launches=1):
    """Algorithmic HBM bytes of answering nq queries in `launches` launches:
    read the queries, write the 8-byte pairs, and one root-to-leaf descent
    per query (32 bytes per level) capped at reading the whole tree (48
    bytes per box) once per launch."""
    descent = nq * math.ceil(math.log2(max(n_boxes, 2))) * 32
    return nq * s_q + pairs * 8 + min(descent, 48 * n_boxes * launches)


def profiled_traffic(workload, algorithmic=None):
    """HBM bytes per launch of the workload's dominant kernel, from the
    committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary of the same
    workload (profiles/traffic.json), or None. With `algorithmic`, the
    profiled traffic / algorithmic-bytes ratio is applied to this run's
    algorithmic bytes per launch (batch composition varies from run to run)."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            t = json.load(f)[workload]
        total = t["fetch_bytes"] + t["write_bytes"]
        if algorithmic is None:
            return int(total), t["source"]
        ratio = total / t["algorithmic_bytes"]
        return int(ratio * algorithmic), f"{t['source']}: traffic {ratio:.2f}x the algorithmic bytes"
    except (OSError, KeyError, ValueError, ZeroDivisionError):
        return None, None


def _median_time(fn, reps):
    """Median wall time of `reps` runs of fn() after one warm-up run; returns
    (seconds, fn's last result)."""
    out = fn()
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        out = fn()
        ts.append(time.perf_counter() - t0)
    return float(np.median(ts)), out


def cpu_baseline(boxes, samples, bf_queries=32, bf_pred=None, reps=3):
    """The oracle's two CPU paths, timed on the host's granted CPUs: (i) the
    STR R-tree (fan-out 16, bulk load timed separately) answering the sampled
    queries, and (ii) the brute-force predicate scan of `bf_queries` of them
    against every box. Counting mode; median of `reps` runs after a warm-up.
    `samples` is a list of (predicate, queries) groups; `value` is the R-tree
    rate, the scan's is reported beside it."""
    from oracle import oracle as O

    ci = cpu_info()
    threads = ci["threads"]
    t_build, tree = _median_time(lambda: O.RTree(boxes, threads=threads), 1)
    nq = sum(qs.shape[0] for _, qs in samples)
    t_rt, npairs = _median_time(lambda: sum(tree.count(p, qs, threads) for p, qs in samples), reps)
    pred0, q0 = samples[-1] if bf_pred is None else next(x for x in samples if x[0] == bf_pred)
    qb = q0[:bf_queries]
    t_bf, n_bf = _median_time(lambda: O.bf_count(pred0, boxes, qb, threads=threads), 1)
    del tree
    return {
        "value": round(nq / t_rt / 1e6, 6),
        "unit": "Mqueries/s",
        "cores": threads,
        "kind": "port",
        "cpu": ci,
        "rtree_mq_per_s": round(nq / t_rt / 1e6, 6),
        "rtree_build_mboxes_per_s": round(boxes.shape[0] / t_build / 1e6, 3),
        "brute_force_mq_per_s": round(qb.shape[0] / t_bf / 1e6, 6),
        "sample": (f"oracle STR R-tree (fan-out 16) over all {boxes.shape[0]} boxes, {nq} queries "
                   f"of the workload ({npairs} results, counted; median of {reps}); bulk load "
                   f"{boxes.shape[0] / t_build / 1e6:.3f} Mboxes/s; brute-force scan of "
                   f"{qb.shape[0]} queries x {boxes.shape[0]} boxes ({n_bf} results): "
                   f"{qb.shape[0] / t_bf / 1e6:.6f} Mqueries/s; {_cores_note(ci)}"),
    }


def cpu_baseline_f64(boxes, pred, queries, threads):
    """The oracle's exact double brute-force join (it has no double R-tree)
    on a sample of the workload's queries."""
    from oracle import oracle as O

    t0 = time.perf_counter()
    n = O.bf_count64(pred, boxes, queries, threads=threads)
    t = time.perf_counter() - t0
    return {
        "value": round(queries.shape[0] / t / 1e6, 6),
        "unit": "Mqueries/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"oracle brute-force double join: {queries.shape[0]} queries drawn from the "
                   f"workload x all {boxes.shape[0]} boxes ({n} results, counted) on "
                   f"{threads} threads"),
    }


def build_rate(R, ix, boxes, torch):
    """Median LBVH build time over 3 rebuilds, phase-timed on the GPU."""
    R.profile_enable(True)
    ms, sort_ms = [], []
    for _ in range(3):
        ix.Clear()
        ix.Insert(boxes)
        torch.cuda.synchronize()
        R.profile_read(reset=True)
        ix.Build()
        ix.Sync()
        p = R.profile_read(reset=True)
        ms.append(sum(p[k][1] for k in ("build_morton", "build_sort", "build_tree")))
        sort_ms.append(p["build_sort"][1])
    R.profile_enable(False)
    n = boxes.shape[0]
    b, s = float(np.median(ms)), float(np.median(sort_ms))
    # Algorithmic bytes: 188 per box for the build (Morton 24, sort 68, leaf
    # gather 32, hierarchy 16, refit 48), 68 per key for the 4-pass sort.
    return {"ms": round(b, 4), "mboxes_per_s": round(n / b / 1e3, 3), "sort_ms": round(s, 4),
            "sort_mkeys_per_s": round(n / s / 1e3, 3),
            "build_hbm_frac": round(188 * n / (b / 1e3) / 1e9 / HBM_PEAK_GBS, 4),
            "sort_hbm_frac": round(68 * n / (s / 1e3) / 1e9 / HBM_PEAK_GBS, 4)}


# ---------------------------------------------------------------- served --
def run_served(args):
    import torch

    import rtspatial_amd as R
    from rtspatial_amd import serve as S
    from rtspatial_amd import workloads as W

    dev = torch.device("cuda:0")
    n_boxes = args.boxes
    log(f"generating the served mixed stream ({n_boxes} boxes, {args.requests} requests)")
    st_gen = W.ServedStream(n_boxes=n_boxes, n_requests=args.requests,
                            max_request=args.max_request)
    boxes_np = st_gen.boxes
    boxes = torch.from_numpy(boxes_np).to(dev)
    env_pool = torch.from_numpy(st_gen.env_pool).to(dev)
    pt_pool = torch.from_numpy(st_gen.pt_pool).to(dev)
    n_req = args.requests
    kinds, sizes = st_gen.kinds, st_gen.sizes

    ix = R.SpatialIndex()
    ix.Init(R.Config(max_geometries=n_boxes))
    build = build_rate(R, ix, boxes, torch)
    log(f"build {build['ms']} ms ({build['mboxes_per_s']} Mboxes/s)")
    tuned = {}
    if not args.no_tune:
        for name, pred, pool in (("point", R.Predicate.kContains, pt_pool),
                                 ("within", R.Predicate.kWithin, env_pool),
                                 ("intersects", R.Predicate.kIntersects, env_pool)):
            t0 = time.perf_counter()
            tuned[name] = ix.Autotune(pred, pool)
            tuned[name]["wall_ms"] = round((time.perf_counter() - t0) * 1e3, 2)
            log(f"autotune {name}: {tuned[name]}")

    preds = [R.Predicate.kContains, R.Predicate.kWithin, R.Predicate.kIntersects]
    slices = []
    for i in range(n_req):
        o, n = int(st_gen.offsets[i]), int(sizes[i])
        q = pt_pool[o:o + n] if kinds[i] == 0 else env_pool[o:o + n]
        slices.append((q, preds[int(kinds[i])]))
    total_q = int(sizes.sum())
    upd = [(torch.from_numpy(ids).to(dev), torch.from_numpy(nb).to(dev))
           for ids, nb in st_gen.updates]

    def trace(rate, with_updates=True):
        t = st_gen.arrivals(rate)
        reqs = [S.Request(queries=q, predicate=p, arrival_s=float(a)) for (q, p), a in zip(slices, t)]
        ups = []
        if with_updates:
            horizon = float(t[-1]) + 1.0 if rate else 5.0
            for i, a in enumerate(st_gen.update_times(horizon)):
                ups.append(S.UpdateBatch(ids=upd[i][0], boxes=upd[i][1], arrival_s=a))
        return reqs, ups

    def reset_index():
        ix.Clear()
        ix.Insert(boxes)
        ix.Build()
        ix.Sync()

    # Served: 3 batches in flight, 6e9-pair arenas, 128K-query chunks, and a
    # request may take 75% of a batch's remaining budget while others wait.
    # Plain: one query call per 256K-query chunk into one 3e9-pair queue.
    cfgs = {"batched": dict(arena_pairs=args.arena_pairs, chunk_queries=128 << 10, slots=3,
                            head_share_pct=75),
            "plain": dict(arena_pairs=3_000_000_000, chunk_queries=256 << 10)}

    # Every batched write launch of this workload is timed (HIP events on its
    # slot's stream) and its algorithmic bytes are summed, so the roofline
    # averages over the same launches as a kernel trace of the whole command.
    acc = {"launches": 0, "ms": 0.0, "bytes": 0}

    def account(prof, queries, s_q_bytes, results):
        wn, wms = prof["batch_write"]
        if wn:
            acc["launches"] += wn
            acc["ms"] += wms
            acc["bytes"] += query_bytes(queries, 0, results, n_boxes, launches=wn) + s_q_bytes

    def one(mode, rate, with_updates=True, **kw):
        reset_index()
        reqs, ups = trace(rate, with_updates)
        R.profile_enable(True)
        R.profile_read(True)
        res, st = S.run(ix, reqs, ups, mode=mode, **{**cfgs[mode], **kw})
        prof = R.profile_read(True)
        R.profile_enable(False)
        nq = sum(int(r.queries.shape[0]) for r in reqs)
        sq = sum(int(r.queries.shape[0]) * (8 if r.queries.shape[1] == 2 else 16) for r in reqs)
        account(prof, nq, sq, int(st["results"]))
        lat = np.array([r["done_s"] - q.arrival_s for r, q in zip(res, reqs)])
        out = {"mode": mode, "rate_rps": rate, "wall_s": round(st["wall_s"], 4),
               "mq_per_s": round(st["queries"] / st["wall_s"] / 1e6, 4),
               "p50_ms": round(float(np.percentile(lat, 50)) * 1e3, 3),
               "p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 3),
               "results": int(st["results"]), "batches": int(st["batches"]),
               "relaunches": int(st["relaunches"]), "updates": int(st["updates"])}
        log(json.dumps(out))
        return out, res, lat

    # Warm-up, then saturation (all requests at t = 0) of both paths.
    for _ in range(args.warmup):
        one("batched", 0.0, with_updates=False)
    sat_plain, _, _ = one("plain", 0.0)
    sat_served, _, _ = one("batched", 0.0)

    # Operating point: 70% of the plain path's capacity (requests per second).
    plain_rps = n_req / sat_plain["wall_s"]
    lam0 = 0.7 * plain_rps
    op_plain, _, _ = one("plain", lam0)
    target = op_plain["p99_ms"]
    # Load sweep of the served path: fixed points, then further points 0.15x
    # apart while p99 holds the target (with a 5% margin, so that the timed
    # replays below hold it too) and the offered rate stays within the served
    # path's saturated rate, and last that rate itself.
    sweep, best = [], None
    served_rps = n_req / sat_served["wall_s"]
    mults = [1.0, 1.5, 2.0, 2.3, 2.6, 2.9]
    at_saturation = False
    i = 0
    while i < len(mults):
        m = mults[i]
        s, _, _ = one("batched", lam0 * m)
        sweep.append(s)
        ok = s["p99_ms"] <= 0.95 * target
        if ok and (best is None or s["mq_per_s"] > best["mq_per_s"]):
            best = s
        if i == len(mults) - 1 and ok and not at_saturation:
            if lam0 * (m + 0.15) <= served_rps:
                mults.append(round(m + 0.15, 2))
            elif lam0 * m < served_rps:
                mults.append(served_rps / lam0)  # last point: the saturated rate itself
                at_saturation = True
        i += 1
    if best is None:
        best = sweep[0]
    op_rate = best["rate_rps"]

    # The timed steps: K replays of the trace at the chosen offered load;
    # value = median throughput over them, latency percentiles over all of
    # their requests.
    log(f"timed: {args.steps} replays at {op_rate:.1f} requests/s")
    steps, lats = [], []
    for _ in range(max(1, args.steps)):
        s, _, lat = one("batched", op_rate)
        steps.append(s)
        lats.append(lat)
    lat_all = np.concatenate(lats)
    mq = np.array([s["mq_per_s"] for s in steps])
    walls = np.array([s["wall_s"] for s in steps])
    p50 = round(float(np.percentile(lat_all, 50)) * 1e3, 3)
    p99 = round(float(np.percentile(lat_all, 99)) * 1e3, 3)
    value = round(float(np.median(mq)), 4)

    # Result check: one more replay at the same load, with every request's
    # pair checksum computed on the device, against the oracle on the index
    # state of each request's epoch (outside the timed region).
    log("verification replay with per-request checksums")
    vrun, vres, _ = one("batched", op_rate, hash_results=1)
    check = {"requests": 0, "oracle_equal": None}
    if not args.no_verify:
        from oracle import served_check as SC
        check = SC.check(st_gen, vres, threads=cpu_threads(), log=log)
        check["replay"] = vrun
        log(f"oracle check: {json.dumps({k: v for k, v in check.items() if k != 'replay'})}")

    # Roofline of the dominant kernel (the batched write pass) over every
    # batched launch of the command. The write kernel alone: one saturated
    # replay with one batch in flight, so no launch shares the GPU.
    log("standalone write-kernel replay (one batch in flight)")
    a0 = dict(acc)
    one("batched", 0.0, with_updates=False, slots=1)
    solo_n = acc["launches"] - a0["launches"]
    solo_ms = acc["ms"] - a0["ms"]
    solo_bytes = acc["bytes"] - a0["bytes"]
    solo_gbs = solo_bytes / (solo_ms / 1e3) / 1e9 if solo_ms > 0 else 0.0
    wn = max(acc["launches"], 1)
    wbytes = acc["bytes"] / wn
    wavg = acc["ms"] / wn
    achieved = wbytes / (wavg / 1e3) / 1e9 if wavg > 0 else 0.0
    traffic, traffic_src = profiled_traffic("served", wbytes)

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "Mqueries/s",
        "n_gpus": 1,
        "steps": len(steps),
        "warmup": args.warmup,
        "ms_per_step": round(float(np.median(walls)) * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": (f"served mixed stream: {n_boxes} Gaussian-mixture boxes (16 components, "
                         f"sigma 0.02, log-normal sides mu=ln 1e-4) resident; {n_req} requests, "
                         f"Poisson arrivals, sizes log-uniform in [1, {args.max_request}] queries "
                         f"({total_q} queries), 40% point / 30% within / 30% intersects with "
                         f"log-normal query sides mu=ln 5e-3; 100k-box update batches at 10 Hz; "
                         f"served = continuous batching (3 batches in flight, persistent "
                         f"work-stealing batches, backfill admission), plain = one query call "
                         f"per request chunk in arrival order; value = median served Mqueries/s "
                         f"over the timed replays at the highest swept offered load whose p99 "
                         f"is within 95% of the plain path's p99 at 70% of its capacity"),
            "path": "served",
            "boxes": n_boxes,
            "requests": n_req,
            "queries": total_q,
            "step": "one replay of the whole request trace at the chosen offered load",
            "seeds": {"boxes": 3001, "queries": 5002, "trace": 5004, "updates": 5003},
        },
        "latency_ms": {"p50": p50, "p99": p99},
        "timed_replays": {"offered_rps": op_rate, "mq_per_s": mq.tolist(),
                          "wall_s": walls.tolist(),
                          "p99_ms_per_replay": [s["p99_ms"] for s in steps]},
        "plain": {"capacity_mq_per_s": sat_plain["mq_per_s"], "at_70pct": op_plain},
        "served": {"capacity_mq_per_s": sat_served["mq_per_s"], "sweep": sweep},
        "p99_target_ms": target,
        "p99_target_met": bool(p99 <= target),
        "served_vs_plain_at_equal_p99": round(value / op_plain["mq_per_s"], 3),
        "served_vs_plain_capacity": round(sat_served["mq_per_s"] / sat_plain["mq_per_s"], 3),
        "bit_exact_check": check,
        "build": build,
        "tuned_plans": tuned,
        "roofline": {
            "bound": "hbm",
            "kernel": "batched write pass of a batch, every batch of the command",
            "achieved": round(achieved, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "traffic_source": traffic_src,
            "algorithmic_bytes_per_launch": int(wbytes),
            "launches": int(wn),
            "avg_launch_ms": round(wavg, 4),
            "note": ("up to 3 batches share the GPU, so launches overlap and their durations "
                     "include the other slots' work; 'standalone' has one batch in flight"),
            "standalone": {"launches": int(solo_n), "avg_launch_ms": round(solo_ms / max(solo_n, 1), 4),
                           "algorithmic_bytes_per_launch": int(solo_bytes / max(solo_n, 1)),
                           "achieved": round(solo_gbs, 2), "frac": round(solo_gbs / HBM_PEAK_GBS, 4)},
        },
    }
    if not args.no_cpu_baseline:
        log("cpu baseline (oracle R-tree / brute force on a sample)")
        rs = np.random.default_rng(7)
        pool_n = st_gen.pool_n
        samples = [(1, st_gen.pt_pool[rs.choice(pool_n, 16000, replace=False)]),
                   (2, st_gen.env_pool[rs.choice(pool_n, 12000, replace=False)]),
                   (0, st_gen.env_pool[rs.choice(pool_n, 12000, replace=False)])]
        out["cpu_baseline"] = cpu_baseline(boxes_np, samples, bf_queries=64)
    else:
        out["cpu_baseline"] = None
    return out


# -------------------------------------------------------------------- c3 --
JOINS = {
    # workload: (description, predicate name, seeds)
    "c1": ("plain point query, small: 100000 boxes (corner U[0,1)^2, sides U[0,0.01]) x "
           "100000 query points U[0,1)^2; Query(kContains, points), one launch per step, all "
           "pairs collected on the device", "kContains", {"boxes": 1001, "queries": 1002}),
    "c2": ("range-contains: 1000000 boxes (sides U[0,1e-3]) x 1000000 query envelopes (sides "
           "U[0,0.01]); every data box inside each query (kWithin), one launch per step, all "
           "pairs collected on the device", "kWithin", {"boxes": 2001, "queries": 2002}),
    "c3": None, "c3f64": None,
}


def run_join(args):
    """One plain Query per step over resident boxes and queries: c1 (points,
    contains), c2 (envelopes, within), c3 / c3f64 (skewed envelopes,
    intersects; c3f64 on a double index)."""
    import torch

    import rtspatial_amd as R
    from rtspatial_amd import workloads as W

    dev = torch.device("cuda:0")
    wl = args.workload
    f64 = wl == "c3f64"
    if wl == "c1":
        boxes_np, queries_np = W.config1()
    elif wl == "c2":
        boxes_np, queries_np = W.config2()
    else:
        boxes_np, queries_np = W.config3(args.boxes, args.queries,
                                         dtype=np.float64 if f64 else np.float32)
    n_boxes = boxes_np.shape[0]
    log(f"{wl}: {n_boxes} boxes, {queries_np.shape[0]} queries, "
        f"{'double' if f64 else 'float'} coordinates")
    boxes = torch.from_numpy(boxes_np).to(dev)
    queries = torch.from_numpy(queries_np).to(dev)
    points = queries_np.shape[1] == 2
    pred = getattr(R.Predicate, JOINS[wl][1]) if JOINS.get(wl) else R.Predicate.kIntersects
    ix = R.SpatialIndex(torch.float64 if f64 else torch.float32)
    ix.Init(R.Config(max_geometries=n_boxes))
    build = build_rate(R, ix, boxes, torch)
    log(f"build: {build}")
    tuned = ix.Autotune(pred, queries) if not args.no_tune else None
    log(f"autotune: {tuned}")
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
    ix.Sync()
    assert queue.size() == n_pairs and not queue.overflowed()
    # Kernel durations: the same steps again with HIP events around each
    # phase (kept out of the timed steps above).
    R.profile_enable(True)
    R.profile_read(reset=True)
    for k in range(args.steps):
        queue.Clear()
        ix.Query(pred, queries, queue)
    torch.cuda.synchronize()
    prof = R.profile_read(reset=True)
    R.profile_enable(False)
    ix.Sync()
    step_ms = np.array([s.elapsed_time(e) for s, e in zip(starts, ends)])
    ms_per_step = wall * 1e3 / args.steps
    nq = queries.shape[0]
    trav_n, trav_ms = prof["traverse"]
    trav_avg = trav_ms / max(trav_n, 1)
    s_q = (8 if points else 16) * (2 if f64 else 1)
    qbytes = query_bytes(nq, s_q, n_pairs, n_boxes)
    achieved = qbytes / (trav_avg / 1e3) / 1e9
    full = wl in ("c1", "c2") or (n_boxes == 10_000_000 and nq == 1_000_000)
    traffic, traffic_src = profiled_traffic(wl) if full else (None, None)
    two_pass = prof["query_count"][0] > 0
    pname = {"c1": "contains, points", "c2": "within"}.get(wl, "intersects")
    kernel = (f"k_query_wave<{pname}, write>" if two_pass else f"k_query_thread<{pname}>")
    if JOINS.get(wl):
        desc, _, seeds = JOINS[wl]
    else:
        desc = (f"range-intersects, skewed: {n_boxes} Gaussian-mixture boxes x {nq} query "
                f"envelopes{' with double coordinates (exact double predicates)' if f64 else ''}; "
                f"plain per-batch LBVH, one Query(kIntersects) per step, all pairs collected on "
                f"the device")
        seeds = {"boxes": 3001, "queries": 3002}
    out = {
        "metric": METRIC,
        "value": round(nq / (ms_per_step / 1e3) / 1e6, 4),
        "unit": "Mqueries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if f64 else "f32",
        "data": "synthetic",
        "config": {"workload": desc, "path": "plain", "boxes": n_boxes, "queries": nq,
                   "pairs_per_step": n_pairs, "seeds": seeds},
        "latency_ms": {"p50": round(float(np.percentile(step_ms, 50)), 4),
                       "p99": round(float(np.percentile(step_ms, 99)), 4)},
        "build": build,
        "tuned_plan": tuned,
        "phases_ms_per_step": {"count_scan": round(prof["query_count"][1] / max(prof["query_count"][0], 1), 4),
                               "order": round(prof["query_order"][1] / max(args.steps, 1), 4),
                               "write": round(trav_avg, 4)},
        "roofline": {"bound": "hbm", "kernel": kernel,
                     "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "traffic_source": traffic_src,
                     "algorithmic_bytes_per_launch": int(qbytes),
                     "avg_launch_ms": round(trav_avg, 5)},
    }
    if not args.no_cpu_baseline:
        log("cpu baseline")
        if f64:
            rs = np.random.default_rng(99)
            out["cpu_baseline"] = cpu_baseline_f64(
                boxes_np, int(pred), queries_np[rs.choice(nq, 1000, replace=False)], cpu_threads())
        else:
            # R-tree over every query of the workload; brute-force scan of
            # every query (c1) or of a fixed sample of 10000 (c2, c3).
            rs = np.random.default_rng(99)
            perm = rs.permutation(nq)
            qs = queries_np[perm]
            out["cpu_baseline"] = cpu_baseline(boxes_np, [(int(pred), qs)],
                                               bf_queries=nq if wl == "c1" else 10_000, reps=3)
    else:
        out["cpu_baseline"] = None
    return out


# ------------------------------------------------------------------ lines --
def run_lines(args):
    """Line-segment intersects (Query(kIntersects, ArrayView<line_t>),
    spatial_index.cuh:517-547): the segments are the diagonals of the skewed
    query envelopes of --workload c3 (min corner to max corner, as the
    reference's GenerateUniformLines draws them, test_commons.h:69-97), over
    the same 10M boxes; all pairs collected on the device. `value` is the
    library's default plan; --sweep also times each traversal mode."""
    import torch

    import rtspatial_amd as R
    from rtspatial_amd import workloads as W

    dev = torch.device("cuda:0")
    log(f"generating workload: {args.boxes} boxes, {args.queries} segments")
    boxes_np, qboxes = W.config3(args.boxes, args.queries)
    segs_np = np.ascontiguousarray(qboxes)  # (x0, y0, x1, y1) read as p1 = min, p2 = max
    boxes = torch.from_numpy(boxes_np).to(dev)
    segs = torch.from_numpy(segs_np).to(dev)
    pred = R.Predicate.kIntersects
    nq = segs.shape[0]
    results = {}
    value_plan = None
    travs = (0, R._native.RTS_TRAVERSAL_THREAD, R._native.RTS_TRAVERSAL_WAVE) if args.sweep else (0,)
    for trav in travs:
        ix = R.SpatialIndex()
        kw = {"leaf_size": args.leaf_size} if args.leaf_size else {}
        ix.Init(R.Config(max_geometries=args.boxes, traversal=trav, **kw))
        build = build_rate(R, ix, boxes, torch)
        cnt = R.SharedValue()
        cnt.set(0)
        ix.QueryLines(pred, segs, cnt)
        n_pairs = cnt.get()
        queue = R.Queue(n_pairs + 1024)
        for _ in range(args.warmup):
            queue.Clear()
            ix.QueryLines(pred, segs, queue)
        torch.cuda.synchronize()
        R.profile_enable(True)
        R.profile_read(reset=True)
        t0 = time.perf_counter()
        for _ in range(args.steps):
            queue.Clear()
            ix.QueryLines(pred, segs, queue)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        prof = R.profile_read(reset=True)
        R.profile_enable(False)
        ix.Sync()
        assert queue.size() == n_pairs and not queue.overflowed()
        trav_n, trav_ms = prof["traverse"]
        name = {0: "default", 1: "thread", 2: "wave"}[trav]
        results[name] = {"ms_per_step": round(wall * 1e3 / args.steps, 4),
                         "traverse_ms": round(trav_ms / max(trav_n, 1), 4),
                         "count_scan_ms": round(prof["query_count"][1] / max(args.steps, 1), 4),
                         "pairs": n_pairs}
        log(f"lines {name}: {results[name]}")
        if trav == 0:
            value_plan = ix
            plan_pairs = n_pairs
            plan_build = build
        else:
            del ix
        del queue
        torch.cuda.empty_cache()
    # results against the oracle on a sample of the segments
    rs = np.random.default_rng(98)
    sample = segs_np[rs.choice(nq, 200, replace=False)]
    c = R.SharedValue()
    c.set(0)
    value_plan.QueryLines(pred, torch.from_numpy(sample).to(dev), c)
    got_sample = c.get()
    from oracle import oracle as O
    t0 = time.perf_counter()
    want_sample = O.bf_count(int(pred), boxes_np, sample, threads=cpu_threads(), lines=True)
    t_cpu = time.perf_counter() - t0
    agree = got_sample == want_sample
    log(f"sample of 200 segments: gpu {got_sample} oracle {want_sample}")
    base = results["default"]
    ms_per_step = base["ms_per_step"]
    qbytes = query_bytes(nq, 16, plan_pairs, args.boxes)
    achieved = qbytes / (base["traverse_ms"] / 1e3) / 1e9
    traffic, traffic_src = profiled_traffic("lines") \
        if args.boxes == 10_000_000 and nq == 1_000_000 else (None, None)
    return {
        "metric": "Mqueries/s, line-segment intersects, 10M-box index (plain LBVH)",
        "value": round(nq / (ms_per_step / 1e3) / 1e6, 4),
        "unit": "Mqueries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": (f"line-segment intersects: {args.boxes} Gaussian-mixture boxes x {nq} "
                         f"segments (the diagonals of the range-intersects query envelopes); "
                         f"one QueryLines(kIntersects) per step, all pairs collected on the device"),
            "path": "plain", "boxes": args.boxes, "queries": nq, "pairs_per_step": plan_pairs,
            "seeds": {"boxes": 3001, "queries": 3002},
        },
        "build": plan_build,
        "traversals": results,
        "results_agree": bool(agree),
        "sample_check": {"segments": 200, "gpu_pairs": got_sample, "oracle_pairs": want_sample},
        "roofline": {"bound": "hbm", "kernel": "k_query_*<segment, write> (the default plan's)",
                     "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "traffic_source": traffic_src,
                     "algorithmic_bytes_per_launch": int(qbytes)},
        "cpu_baseline": None if args.no_cpu_baseline else {
            "value": round(sample.shape[0] / t_cpu / 1e6, 6),
            "unit": "Mqueries/s",
            "cores": cpu_threads(),
            "kind": "port",
            "sample": (f"oracle brute-force exact segment join: 200 segments drawn from the "
                       f"workload x all {args.boxes} boxes ({want_sample} results, counted) on "
                       f"{cpu_threads()} threads"),
        },
    }


# -------------------------------------------------------------- handlers --
def run_handlers(args):
    """The 10M x 1M intersects join with result handlers compiled outside the
    library (rtspatial_amd/examples/librts_user_handlers.so, built on the
    public device headers), against the built-in handlers on the same index."""
    import ctypes

    import torch

    import rtspatial_amd as R
    from rtspatial_amd import build as B
    from rtspatial_amd import workloads as W

    dev = torch.device("cuda:0")
    boxes_np, queries_np = W.config3(args.boxes, args.queries)
    boxes = torch.from_numpy(boxes_np).to(dev)
    queries = torch.from_numpy(queries_np).to(dev)
    nq = queries.shape[0]
    pred = R.Predicate.kIntersects
    ix = R.SpatialIndex()
    ix.Init(R.Config(max_geometries=args.boxes))
    ix.Insert(boxes)
    ix.Build()
    tuned = ix.Autotune(pred, queries) if not args.no_tune else None
    log(f"autotune: {tuned}")
    R._native.lib()
    X = ctypes.CDLL(B.USER_HANDLERS_LIB)
    vp, u64, i32, u32 = ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int, ctypes.c_uint32
    X.rtsx_summary.argtypes = [vp, i32, vp, u64, i32, i32, u32, vp, vp, vp, vp]
    X.rtsx_count.argtypes = [vp, i32, vp, u64, i32, i32, u32, vp, vp]
    X.rtsx_reference_count.argtypes = [vp, i32, vp, u64, i32, i32, vp, vp]
    X.rtsx_last_error.restype = ctypes.c_char_p
    stream = torch.cuda.current_stream()
    sp = ctypes.c_void_p(stream.cuda_stream)
    count = torch.zeros(nq, dtype=torch.int64, device=dev)
    hsum = torch.zeros(nq, dtype=torch.int64, device=dev)
    mn = torch.zeros(nq, dtype=torch.int32, device=dev)
    counter = torch.zeros(1, dtype=torch.int64, device=dev)
    cnt = R.SharedValue()
    cnt.set(0)
    ix.Query(pred, queries, cnt)
    n_pairs = cnt.get()
    queue = R.Queue(n_pairs + 1024)

    def builtin_count():
        cnt.tensor.zero_()
        ix.Query(pred, queries, cnt)

    def builtin_collect():
        queue.Clear()
        ix.Query(pred, queries, queue)

    def ok(rc):
        if rc != 0:
            raise RuntimeError(X.rtsx_last_error().decode())

    def summary():
        count.zero_()
        hsum.zero_()
        mn.fill_(-1)
        ok(X.rtsx_summary(ix._handle(), int(pred), queries.data_ptr(), nq, 0, 0, 0,
                          count.data_ptr(), hsum.data_ptr(), mn.data_ptr(), sp))

    def count_done():
        ok(X.rtsx_count(ix._handle(), int(pred), queries.data_ptr(), nq, 0, 0, 0,
                        count.data_ptr(), sp))

    def reference_count():
        counter.zero_()
        ok(X.rtsx_reference_count(ix._handle(), int(pred), queries.data_ptr(), nq, 0, 0,
                                  counter.data_ptr(), sp))

    def timed(fn):
        for _ in range(args.warmup):
            fn()
        torch.cuda.synchronize()
        ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
              for _ in range(args.steps)]
        for a, b in ev:
            a.record()
            fn()
            b.record()
        torch.cuda.synchronize()
        ms = float(np.median([a.elapsed_time(b) for a, b in ev]))
        return {"ms": round(ms, 3), "mq_per_s": round(nq / ms / 1e3, 3)}

    res = {}
    for name, fn in (("builtin_count", builtin_count), ("builtin_collect", builtin_collect),
                     ("count_done", count_done), ("summary", summary),
                     ("reference_count", reference_count)):
        log(f"timing {name}")
        res[name] = timed(fn)
        log(f"{name}: {res[name]}")
    # The handlers' results agree with the built-in ones on the same index.
    count_done()
    torch.cuda.synchronize()
    per_q_done = count.clone()
    summary()
    torch.cuda.synchronize()
    checks = {"summary_total": int(count.sum().item()), "count_done_total": int(per_q_done.sum().item()),
              "per_query_equal": bool(torch.equal(per_q_done, count))}
    reference_count()
    torch.cuda.synchronize()
    checks["reference_total"] = int(counter.item())
    checks["builtin_total"] = int(n_pairs)
    agree = (checks["summary_total"] == checks["count_done_total"] == checks["reference_total"]
             == n_pairs) and checks["per_query_equal"]
    ix.Sync()
    out = {
        "metric": METRIC,
        "value": res["summary"]["mq_per_s"],
        "unit": "Mqueries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": res["summary"]["ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": (f"range-intersects, skewed: {args.boxes} Gaussian-mixture boxes x {nq} "
                         f"query envelopes ({n_pairs} pairs); value = the caller-compiled per-pair "
                         f"summary handler (per query: count, splitmix64 hash sum, min id)"),
            "path": "user handlers", "boxes": args.boxes, "queries": nq, "pairs": n_pairs,
        },
        "handlers": res,
        "checks": checks,
        "results_agree": agree,
        "tuned_plan": tuned,
        "cpu_baseline": None,
    }
    if not agree:
        log(f"MISMATCH between handler results: {checks}")
    return out


# --------------------------------------------------------------- inserts --
def run_inserts(args):
    """Insert batches of growing size into the resident 10M-box index, each
    followed by a query batch: the first query after an Insert pays the build
    (of the small delta tree while the inserts stay within insert_delta_pct of
    the main tree, else of the whole tree), the second does not. The
    reference builds one GAS per Insert under an IAS instead
    (spatial_index.cuh:142-196)."""
    import torch

    import rtspatial_amd as R
    from rtspatial_amd import workloads as W

    dev = torch.device("cuda:0")
    rng = W._rng(3, 1)
    means = W.mixture_means(rng)
    boxes = torch.from_numpy(W.mixture_boxes(rng, means, args.boxes, math.log(1e-4))).to(dev)
    extra = torch.from_numpy(W.mixture_boxes(W._rng(3, 7), means, 1_111_000, math.log(1e-4))).to(dev)
    pts = torch.from_numpy(W.mixture_boxes(W._rng(3, 8), means, 100_000, math.log(1e-4))[:, :2]
                           .copy()).to(dev)
    ix = R.SpatialIndex()
    ix.Init(R.Config(max_geometries=args.boxes + 1_111_000))
    ix.Insert(boxes)
    ix.Build()
    cnt = R.SharedValue()

    def query():
        cnt.tensor.zero_()
        ix.Query(R.Predicate.kContains, pts, cnt)

    def timed(fn):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        a.record()
        fn()
        b.record()
        torch.cuda.synchronize()
        return a.elapsed_time(b)

    query()
    rounds = []
    off = 0
    for batch in (1_000, 10_000, 100_000, 1_000_000):
        part = extra[off:off + batch]
        off += batch
        t_ins = timed(lambda: ix.Insert(part))
        t_first = timed(query)   # rebuild + query
        t_second = timed(query)  # query alone
        rounds.append({"batch": batch, "boxes_after": ix.size(), "insert_ms": round(t_ins, 3),
                       "first_query_ms": round(t_first, 3), "query_ms": round(t_second, 3),
                       "rebuild_ms": round(t_first - t_second, 3),
                       "results": cnt.get()})
        log(json.dumps(rounds[-1]))
    ix.Sync()
    ins = sum(r["batch"] for r in rounds)
    ms = sum(r["insert_ms"] + max(r["rebuild_ms"], 0.0) for r in rounds)
    return {
        "metric": METRIC,
        "value": round(ins / (ms / 1e3) / 1e6, 3),
        "unit": "Mboxes/s inserted (insert + build, all batches)",
        "n_gpus": 1, "steps": len(rounds), "warmup": 0,
        "ms_per_step": round(ms / len(rounds), 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": (f"insert batches of 1e3..1e6 Gaussian-mixture boxes into a resident "
                                f"{args.boxes}-box index, each followed by 100k point queries"),
                   "path": "inserts", "boxes": args.boxes},
        "rounds": rounds,
        "cpu_baseline": None,
    }


# -------------------------------------------------------------------- c4 --
def run_c4(args):
    import torch

    import rtspatial_amd as R
    from rtspatial_amd import workloads as W

    dev = torch.device("cuda:0")
    n_rounds = args.rounds
    log(f"generating mutable-index workload: {args.boxes} boxes, {n_rounds} rounds")
    m = W.MutationRounds(args.boxes, args.boxes // 10, args.queries)
    initial = torch.from_numpy(m.initial).to(dev)
    rounds = []
    for _ in range(n_rounds):
        d, i, u, up, p = m.next_round()
        rounds.append(tuple(torch.from_numpy(np.ascontiguousarray(x)).to(dev) for x in (d, i, u, up, p)))
    final_boxes, final_alive = m.boxes, m.alive
    del m
    policies = {"refit": R._native.RTS_REBUILD_NEVER, "sah": R._native.RTS_REBUILD_SAH,
                "rebuild": R._native.RTS_REBUILD_ALWAYS}
    res, counts = {}, {}
    for name, pol in policies.items():
        for inserts in (True, False):
            ix = R.SpatialIndex()
            ix.Init(R.Config(max_geometries=args.boxes, rebuild_policy=pol))
            ix.Insert(initial)
            ix.Build()
            torch.cuda.synchronize()
            R.profile_enable(True)
            R.profile_read(True)
            cnt = R.SharedValue()
            ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
                  for _ in range(n_rounds)]
            per = []
            t0 = time.perf_counter()
            for k, (d, i, u, up, p) in enumerate(rounds):
                ev[k][0].record()
                ix.Delete(d)
                if inserts:
                    ix.Insert(i)
                ix.Update((u, up))
                cnt.set(0)
                ix.Query(R.Predicate.kContains, p, cnt)
                ev[k][1].record()
                if inserts:
                    per.append(cnt.tensor.clone())
            torch.cuda.synchronize()
            wall = time.perf_counter() - t0
            prof = R.profile_read(True)
            R.profile_enable(False)
            st = ix.stats()
            ms = [a.elapsed_time(b) for a, b in ev]
            key = f"{name}{'' if inserts else '_no_inserts'}"
            res[key] = {"round_ms_median": round(float(np.median(ms)), 3),
                        "round_ms_max": round(float(np.max(ms)), 3), "wall_s": round(wall, 3),
                        "builds": int(st["builds"]), "refits": int(st["refits"]),
                        "phases_ms": {k: round(v[1], 3) for k, v in prof.items() if v[0]}}
            log(f"c4 {key}: {res[key]}")
            if inserts:
                counts[name] = [int(c.item()) for c in per]
            del ix
            torch.cuda.empty_cache()
    same = counts["refit"] == counts["sah"] == counts["rebuild"]
    assert same, "policies disagree on per-round result counts"
    # Final round against the oracle on a sample of the points.
    from oracle import oracle as O
    pts = rounds[-1][4].cpu().numpy()
    sample = pts[np.random.default_rng(11).choice(pts.shape[0], 20000, replace=False)]
    ix = R.SpatialIndex()
    ix.Init(R.Config(max_geometries=args.boxes))
    ix.Insert(initial)
    for d, i, u, up, _ in rounds:
        ix.Delete(d)
        ix.Insert(i)
        ix.Update((u, up))
    cnt = R.SharedValue()
    cnt.set(0)
    ix.Query(R.Predicate.kContains, torch.from_numpy(sample).to(dev), cnt)
    want = O.RTree(final_boxes, alive=final_alive.astype(np.uint8)).count(O.CONTAINS, sample, cpu_threads())
    assert cnt.get() == want, "final round differs from the oracle"
    nq = args.queries
    best = min(("refit", "sah", "rebuild"), key=lambda k: res[k]["round_ms_median"])
    ms = res[best]["round_ms_median"]
    return {
        "metric": "mutable-index round time (1 delete + 1 insert + 1 update batch + 1 point query batch)",
        "value": round(nq / (ms / 1e3) / 1e6, 4),
        "unit": "Mqueries/s (point queries incl. the round's mutations)",
        "n_gpus": 1, "steps": n_rounds, "warmup": 0, "ms_per_step": ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": (f"mutable index: {args.boxes} uniform boxes (sides U[0,1e-3]); "
                                f"{n_rounds} rounds of {args.boxes // 10} deletes, inserts and "
                                f"updates (shift U[-1e-3,1e-3]) then {nq} uniform point queries"),
                   "best_policy": best},
        "policies": res,
        "per_round_counts_equal_across_policies": same,
        "oracle_check": {"sample_points": int(sample.shape[0]), "count": int(want)},
    }


def run_copy_calibration(args):
    """A plain 1 GiB device-to-device copy, three times: the known-size
    kernel whose FETCH_SIZE / WRITE_SIZE readings scale the counters of the
    index's kernels to bytes."""
    import torch

    n = 1 << 30
    src = torch.empty(n, dtype=torch.uint8, device="cuda:0").fill_(1)
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / 3
    return {"metric": "device copy calibration", "value": round(2 * n / dt / 1e9, 2), "unit": "GB/s",
            "n_gpus": 1, "bytes_read_per_copy": n, "bytes_written_per_copy": n}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload",
                    choices=["served", "c1", "c2", "c3", "c3f64", "lines", "handlers", "inserts",
                             "c4", "copy"],
                    default="served")
    ap.add_argument("--rounds", type=int, default=20)
    ap.add_argument("--boxes", type=int, default=10_000_000)
    ap.add_argument("--queries", type=int, default=1_000_000)
    ap.add_argument("--requests", type=int, default=600)
    ap.add_argument("--max-request", type=int, default=1_000_000)
    ap.add_argument("--arena-pairs", type=int, default=6_000_000_000)
    ap.add_argument("--no-verify", action="store_true",
                    help="served: skip the oracle check of the verification replay")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-tune", action="store_true", help="skip the autotuner (default plans)")
    ap.add_argument("--sweep", action="store_true", help="lines: also time each traversal mode")
    ap.add_argument("--leaf-size", type=int, default=0,
                    help="lines: leaf-range size of the index (0: the library default)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "bench.py measures one GPU; run with --gpus 1",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    import torch

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    torch.cuda.set_device(0)
    out = {"served": run_served, "c1": run_join, "c2": run_join, "c3": run_join,
           "c3f64": run_join, "lines": run_lines,
           "handlers": run_handlers,
           "inserts": run_inserts, "c4": run_c4,
           "copy": run_copy_calibration}[args.workload](args)
    print(json.dumps(out), flush=True)
    check = out.get("bit_exact_check")
    if check and check.get("oracle_equal") is False:
        sys.exit(3)  # results differ from the oracle: the line above says where
    if out.get("results_agree") is False:
        sys.exit(3)


if __name__ == "__main__":
    main()
