"""WKT ingest in Python, with the semantics of include/rtspatial/io/wkt_loader.h
(and of the reference's examples/wkt_loader.h:25-129):

  load_boxes(path, limit)   float32 [n, 4] (min x, min y, max x, max y): one box
                            per POLYGON line, one per member of a MULTIPOLYGON,
                            the min/max over each outer ring's vertices.
  load_points(path, limit)  float32 [n, 2]: POINT lines, plus every outer-ring
                            vertex of POLYGON / MULTIPOLYGON lines.

Coordinates are rounded to float32 directly from their decimal text (C strtof),
as the C++ loader does; reading stops after the line on which the count
reaches `limit`. Any other geometry raises ValueError ("Bad Geometry").
"""
import ctypes
import re

import numpy as np

_libc = ctypes.CDLL(None)
_libc.strtof.restype = ctypes.c_float
_libc.strtof.argtypes = [ctypes.c_char_p, ctypes.c_void_p]

_NUM = re.compile(r"[-+]?(?:\d+\.?\d*|\.\d+)(?:[eE][-+]?\d+)?")


def _f32(tok):
    return _libc.strtof(tok.encode(), None)


def _outer_rings(line):
    if line.startswith("MULTIPOLYGON"):
        body = line[len("MULTIPOLYGON"):].strip()
        if body.startswith("EMPTY"):
            return []
        # MULTIPOLYGON(((outer), (hole)), ((outer)))
        return _polygons(body, level=2)
    if line.startswith("POLYGON"):
        body = line[len("POLYGON"):].strip()
        if body.startswith("EMPTY"):
            return []
        return _polygons(body, level=1)
    raise ValueError(f"Bad Geometry {line[:120]}")


def _polygons(body, level):
    """Outer rings: the first ring opened at depth level + 1 inside each
    polygon (depth `level` for a POLYGON body, `level` = 2 in a MULTIPOLYGON)."""
    rings, depth, start, first_in_poly = [], 0, None, True
    for i, ch in enumerate(body):
        if ch == "(":
            depth += 1
            if depth == level:
                first_in_poly = True
            elif depth == level + 1 and first_in_poly:
                start = i + 1
        elif ch == ")":
            if depth == level + 1 and start is not None:
                nums = [_f32(t) for t in _NUM.findall(body[start:i])]
                if len(nums) < 2:
                    raise ValueError(f"Bad Geometry (empty ring) {body[:120]}")
                rings.append((nums[0::2], nums[1::2]))
                start = None
                first_in_poly = False
            depth -= 1
    if depth != 0:
        raise ValueError(f"Bad Geometry (unbalanced parentheses) {body[:120]}")
    return rings


def load_boxes(path, limit=None):
    out = []
    with open(path) as f:
        for line in f:
            line = line.strip()
            if not line:
                continue
            for xs, ys in _outer_rings(line):
                out.append((min(xs), min(ys), max(xs), max(ys)))
            if limit is not None and len(out) >= limit:
                break
    return np.asarray(out, dtype=np.float32).reshape(-1, 4)


def load_points(path, limit=None):
    out = []
    with open(path) as f:
        for line in f:
            line = line.strip()
            if not line:
                continue
            if line.startswith("POINT"):
                nums = [_f32(t) for t in _NUM.findall(line[len("POINT"):])]
                if len(nums) < 2:
                    raise ValueError(f"Bad Geometry {line[:120]}")
                out.append(nums[:2])
            else:
                for xs, ys in _outer_rings(line):
                    out.extend(zip(xs, ys))
            if limit is not None and len(out) >= limit:
                break
    return np.asarray(out, dtype=np.float32).reshape(-1, 2)
