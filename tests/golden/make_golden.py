"""Writes the golden fixtures under tests/golden/ (run from the repo root:
`python tests/golden/make_golden.py`; needs the CPU oracle built and the
reference's dataset files, which it reads as data only).

  wkt_1000.npz   inputs parsed from the reference's sample datasets
                 (dataset/box/{uniform,gaussian}_1000.wkt,
                 dataset/point/{uniform,gaussian}_1000.wkt) the way its loader
                 does it (examples/wkt_loader.h:25-129: coordinates parsed as
                 float, box = min/max over the outer ring), plus the exact
                 sorted (query_id, box_id) pair sets of every predicate over
                 them, computed by the oracle's brute force and checked against
                 its R-tree.
  known_answers.json  the reference tests' assertions (test/*.h) with the
                 inputs that are small enough to store, and the one pair by
                 which the reference's intersects count exceeds the exact one.
"""
import ctypes
import json
import os
import re
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from oracle import oracle as O  # noqa: E402

REF = "/root/reference/dataset"
OUT = os.path.dirname(os.path.abspath(__file__))

_libc = ctypes.CDLL(None)
_libc.strtof.restype = ctypes.c_float
_libc.strtof.argtypes = [ctypes.c_char_p, ctypes.c_void_p]


def f32(tok):
    return np.float32(_libc.strtof(tok.encode(), None))


_NUM = re.compile(r"[-+0-9.eE]+")


def parse_boxes(path):
    out = []
    with open(path) as f:
        for line in f:
            line = line.strip()
            if not line:
                continue
            assert line.startswith("POLYGON"), line[:20]
            nums = [f32(t) for t in _NUM.findall(line[len("POLYGON"):])]
            xs, ys = nums[0::2], nums[1::2]
            out.append([min(xs), min(ys), max(xs), max(ys)])
    return np.asarray(out, dtype=np.float32)


def parse_points(path):
    out = []
    with open(path) as f:
        for line in f:
            line = line.strip()
            if not line:
                continue
            assert line.startswith("POINT"), line[:20]
            nums = [f32(t) for t in _NUM.findall(line[len("POINT"):])]
            out.append(nums[:2])
    return np.asarray(out, dtype=np.float32)


def main():
    data = {}
    for kind in ("uniform", "gaussian"):
        data[f"{kind}_boxes"] = parse_boxes(f"{REF}/box/{kind}_1000.wkt")
        data[f"{kind}_points"] = parse_points(f"{REF}/point/{kind}_1000.wkt")
    for bk in ("uniform", "gaussian"):
        boxes = data[f"{bk}_boxes"]
        rt = O.RTree(boxes)
        for qk in ("uniform", "gaussian"):
            pts = data[f"{qk}_points"]
            p = O.bf_pairs(O.CONTAINS, boxes, pts)
            assert np.array_equal(p, rt.pairs(O.CONTAINS, pts))
            data[f"pairs_{bk}_boxes_{qk}_points_contains"] = p
            qb = data[f"{qk}_boxes"]
            for name, pred in (("intersects", O.INTERSECTS), ("contains", O.CONTAINS),
                               ("within", O.WITHIN)):
                p = O.bf_pairs(pred, boxes, qb)
                assert np.array_equal(p, rt.pairs(pred, qb))
                data[f"pairs_{bk}_boxes_{qk}_boxes_{name}"] = p
    np.savez_compressed(os.path.join(OUT, "wkt_1000.npz"), **data)

    # Known answers of the reference's tests.
    b100k = O.ref_uniform_boxes(100000, 0.5, 0.5)
    p1k = O.ref_uniform_points(1000)
    q1k = O.ref_uniform_boxes(1000, 0.1, 0.1)
    ref_count, extra, missing = O.refray_intersects(b100k, q1k)
    upd_ids = O.ref_update_ids(100000, 5000)
    upd_boxes = O.ref_uniform_boxes(5000, 0.5, 0.5)
    b_upd = b100k.copy()
    b_upd[upd_ids.astype(np.int64)] = upd_boxes
    ka = {
        "point_contains_100k_1k": {
            "source": "test/test_point_queries.h:9-60",
            "reference": 4224111,
            "exact": O.bf_count(O.CONTAINS, b100k, p1k),
        },
        "intersects_100k_1k": {
            "source": "test/test_envelope_queries.h:38-91",
            "reference": 6549771,
            "exact": O.bf_count(O.INTERSECTS, b100k, q1k),
            "reference_formulation": ref_count,
            "extra_pairs_query_box": extra.tolist(),
            "missing_pairs_query_box": missing.tolist(),
        },
        "intersects_after_update_5pct": {
            "source": "test/test_envelope_queries.h:93-164",
            "exact": O.bf_count(O.INTERSECTS, b_upd, q1k),
            "update_ids_first5": upd_ids[:5].tolist(),
        },
        "tiny_points": {
            "source": "test/test_point_queries.h:62-120",
            "boxes": [[0, 0, 1, 1], [0.2, 0.2, 0.8, 0.8]],
            "points": [[0.1, 0.1], [0.5, 0.5], [0.5, 0.6], [1.0, 1.1], [2.0, 2.0]],
            "reference": 5,
            "pairs_query_box": [[0, 0], [1, 0], [1, 1], [2, 0], [2, 1]],
        },
        "tiny_intersects_delete": {
            "source": "test/test_envelope_queries.h:11-36,166-245",
            "boxes": [[0, 0, 1, 1], [0.5, -0.5, 1.5, 0.5]],
            "queries": [[0.75, 0.25, 1.25, 0.75]],
            "reference_before": 2,
            "reference_after_delete_1": 1,
        },
    }
    with open(os.path.join(OUT, "known_answers.json"), "w") as f:
        json.dump(ka, f, indent=1)
    print("wrote", sorted(os.listdir(OUT)))


if __name__ == "__main__":
    main()
