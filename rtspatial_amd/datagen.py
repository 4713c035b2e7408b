"""Synthetic spatial datasets in the style of the reference's Spider generator
(dataset/generator.py; its sample files are made by dataset/gen_data.sh).

Distributions (dataset/generator.py:300-460), all in the unit square:
  uniform                      every coordinate U[0, 1)
  gaussian                     every coordinate N(0.5, 0.1)
  diagonal  percentage, buffer  a share on the diagonal, the rest N(0, buffer/5)
                               off it
  sierpinski                   the chaos game on the triangle (0,0) (1,0) (0.5, sqrt3/2)
  bit       probability, digits  coordinates sum_i Bernoulli(p) / 2^i
  parcel    srange, dither     recursive splits of the unit square into
                               `cardinality` boxes, then shrunk by up to `dither`
Points outside [0, 1] are redrawn (generator.py:266-272). With geometry=box, a
point becomes the box point -+ U[0, maxsize] per dimension (PointToBoxSink,
generator.py:150-166); parcel always makes boxes. Output formats: wkt (POINT /
five-vertex POLYGON lines, generator.py:68-94) and csv.

The draws follow the same formulas as the reference (uniform = r*(max-min)+min,
normal by the Box-Muller sine form, dice = floor(r*n)+1) on a random.Random
seeded by `seed` (the reference draws from Python's global generator,
unseeded).

  python -m rtspatial_amd.datagen distribution=uniform cardinality=1000 \\
      dimensions=2 geometry=box maxsize=0.005,0.005 format=wkt seed=1 > boxes.wkt
"""
import math
import random
import sys


class Draw:
    def __init__(self, seed=None):
        self.r = random.Random(seed)

    def uniform(self, lo, hi):
        return self.r.random() * (hi - lo) + lo

    def bernoulli(self, p):
        return 1 if self.r.random() < p else 0

    def normal(self, mu, sigma):
        return mu + sigma * math.sqrt(-2 * math.log(self.r.random())) * math.sin(
            2 * math.pi * self.r.random())

    def dice(self, n):
        return math.floor(self.r.random() * n) + 1


def _valid(p):
    return all(0 <= x <= 1 for x in p)


def points(distribution, card, dim, draw, **kw):
    """Yields `card` points of the distribution (redrawing invalid ones)."""
    prev, i = None, 0
    tri = [[0.0, 0.0], [1.0, 0.0], [0.5, math.sqrt(3) / 2]]
    while i < card:
        if distribution == "uniform":
            p = [draw.r.random() for _ in range(dim)]
        elif distribution == "gaussian":
            p = [draw.normal(0.5, 0.1) for _ in range(dim)]
        elif distribution == "diagonal":
            if draw.bernoulli(kw["percentage"]) == 1:
                p = [draw.r.random()] * dim
            else:
                c = draw.r.random()
                d = draw.normal(0, kw["buffer"] / 5)
                p = [(c + (1 - 2 * (x % 2)) * d / math.sqrt(2)) for x in range(dim)]
        elif distribution == "sierpinski":
            if i < 3:
                p = list(tri[i])
            else:
                d = draw.dice(5)
                corner = tri[0] if d <= 2 else tri[1] if d <= 4 else tri[2]
                p = [(prev[k] + corner[k]) / 2 for k in range(len(prev))]
        elif distribution == "bit":
            p = [sum(draw.bernoulli(kw["probability"]) / math.pow(2, k)
                     for k in range(1, kw["digits"] + 1)) for _ in range(dim)]
        else:
            raise ValueError(f"unknown point distribution {distribution}")
        if _valid(p):
            yield p
            prev = p
            i += 1


def parcel_boxes(card, draw, srange, dither):
    """generator.py:373-432: depth-first splits of the unit square; the last
    level is split only as often as needed to reach `card` boxes."""
    max_height = math.ceil(math.log(card, 2)) if card > 1 else 0
    to_split = card - pow(2, max(max_height - 1, 0))
    stack = [(0, 0.0, 0.0, 1.0, 1.0)]
    split_done = made = 0

    def split(b):
        depth, x, y, w, h = b
        if w > h:
            s = w * draw.uniform(srange, 1 - srange)
            return (depth + 1, x, y, s, h), (depth + 1, x + s, y, w - s, h)
        s = h * draw.uniform(srange, 1 - srange)
        return (depth + 1, x, y, w, s), (depth + 1, x, y + s, w, h - s)

    def dithered(b):
        _, x, y, w, h = b
        dx = w * draw.uniform(0.0, dither)
        x, w = x + dx / 2, w - dx
        dy = h * draw.uniform(0.0, dither)
        y, h = y + dy / 2, h - dy
        return [x, y], [x + w, y + h]

    while made < card:
        b = stack.pop()
        if b[0] >= max_height - 1:
            if split_done < to_split:
                b1, b2 = split(b)
                split_done += 1
                yield dithered(b1)
                yield dithered(b2)
                made += 2
            else:
                yield dithered(b)
                made += 1
        else:
            b1, b2 = split(b)
            stack.append(b2)
            stack.append(b1)


def boxes(distribution, card, dim, draw, maxsize=None, **kw):
    """Yields (min corner, max corner) pairs."""
    if distribution == "parcel":
        yield from parcel_boxes(card, draw, kw["srange"], kw["dither"])
        return
    for p in points(distribution, card, dim, draw, **kw):
        lo, hi = [], []
        for d in range(len(p)):
            size = draw.uniform(0, maxsize[d])
            lo.append(p[d] - size)
            hi.append(p[d] + size)
        yield lo, hi


def wkt_point(p):
    return "POINT(" + " ".join(str(e) for e in p) + ")"


def wkt_box(lo, hi):
    return (f"POLYGON(({lo[0]} {lo[1]},{hi[0]} {lo[1]},{hi[0]} {hi[1]},{lo[0]} {hi[1]},"
            f"{lo[0]} {lo[1]}))")


def csv_point(p):
    return ",".join(str(e) for e in p)


def csv_box(lo, hi):
    return ",".join(str(e) for e in lo) + "," + ",".join(str(e) for e in hi)


def generate(out, distribution, cardinality, dimensions=2, geometry="point", fmt="wkt",
             seed=None, maxsize=None, **kw):
    """Writes the dataset to the text stream `out`, one geometry per line."""
    draw = Draw(seed)
    if geometry == "box" or distribution == "parcel":
        if distribution != "parcel" and maxsize is None:
            raise ValueError("geometry=box needs maxsize")
        enc = wkt_box if fmt == "wkt" else csv_box
        for lo, hi in boxes(distribution, cardinality, dimensions, draw, maxsize=maxsize, **kw):
            out.write(enc(lo, hi) + "\n")
    else:
        enc = wkt_point if fmt == "wkt" else csv_point
        for p in points(distribution, cardinality, dimensions, draw, **kw):
            out.write(enc(p) + "\n")
    out.flush()


def main(argv):
    args = dict(a.split("=", 1) for a in argv[1:] if "=" in a)
    if "distribution" not in args or "cardinality" not in args:
        sys.stderr.write(__doc__)
        return 1
    kw = {}
    for k, conv in (("percentage", float), ("buffer", float), ("probability", float),
                    ("digits", int), ("srange", float), ("dither", float)):
        if k in args:
            kw[k] = conv(args[k])
    maxsize = [float(v) for v in args["maxsize"].split(",")] if "maxsize" in args else None
    generate(sys.stdout, args["distribution"], int(args["cardinality"]),
             int(args.get("dimensions", 2)), args.get("geometry", "point"),
             args.get("format", "csv").lower(), int(args["seed"]) if "seed" in args else None,
             maxsize, **kw)
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv))
