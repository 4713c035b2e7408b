"""Depth bound of the LBVH hierarchy, on the CPU.

The build (build.hip, k_karras) emits a binary radix tree (Karras 2012) over
the sorted 30-bit Morton keys, with equal keys told apart by their positions:
the common prefix of positions i and j is clz(key_i ^ key_j) for different
keys and 32 + clz(i ^ j) for equal ones. Every internal node's prefix length
is strictly greater than its parent's, and the lengths lie in [2, 31] (30-bit
keys) or [34, 63] (positions < 2^30), so no root-to-leaf path has more than
60 internal nodes. The per-lane traversal stack (kStackDepth) must hold
depth + 1 entries. This test restates the split rule in numpy, builds the
deepest inputs the keys allow (a chain of one-bit keys, each level splitting
one key off, above a run of equal keys whose positions add a balanced
subtree) and checks the bound and the stack size."""
import os
import re

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _clz32(x):
    x = np.asarray(x, dtype=np.uint64) & np.uint64(0xFFFFFFFF)
    out = np.full(x.shape, 32, dtype=np.int64)
    nz = x != 0
    out[nz] = 31 - np.floor(np.log2(x[nz].astype(np.float64))).astype(np.int64)
    return out


def _delta(keys, i, j):
    """Common prefix length of (key, position) at sorted positions i and j; -1
    outside the array (k_karras' prefix_len)."""
    n = keys.shape[0]
    i = np.asarray(i, np.int64)
    j = np.asarray(j, np.int64)
    out = np.full(np.broadcast(i, j).shape, -1, np.int64)
    ok = (j >= 0) & (j < n)
    ii, jj = np.broadcast_to(i, out.shape)[ok], np.broadcast_to(j, out.shape)[ok]
    a, b = keys[ii], keys[jj]
    same = a == b
    v = np.where(same, 32 + _clz32(ii ^ jj), _clz32(a ^ b))
    out[ok] = v
    return out


def karras(keys):
    """Children of every internal node (leaf children flagged by >= n - 1 +
    offset), vectorised over all nodes, as k_karras computes them."""
    n = keys.shape[0]
    i = np.arange(n - 1, dtype=np.int64)
    d = np.where(_delta(keys, i, i + 1) - _delta(keys, i, i - 1) >= 0, 1, -1)
    dmin = _delta(keys, i, i - d)
    lmax = np.full(n - 1, 2, np.int64)
    while True:
        grow = _delta(keys, i, i + lmax * d) > dmin
        if not grow.any():
            break
        lmax = np.where(grow, lmax * 2, lmax)
    l = np.zeros(n - 1, np.int64)
    t = lmax // 2
    while (t >= 1).any():
        step = (t >= 1) & (_delta(keys, i, i + (l + t) * d) > dmin)
        l = np.where(step, l + t, l)
        t = t // 2
    j = i + l * d
    dnode = _delta(keys, i, j)
    s = np.zeros(n - 1, np.int64)
    t = l.copy()
    while True:
        t = (t + 1) // 2
        step = _delta(keys, i, i + (s + t) * d) > dnode
        s = np.where(step, s + t, s)
        if (t <= 1).all():
            break
    gamma = i + s * d + np.where(d < 0, -1, 0)
    first, last = np.minimum(i, j), np.maximum(i, j)
    left = np.where(first == gamma, (n - 1) + gamma, gamma)
    right = np.where(last == gamma + 1, (n - 1) + gamma + 1, gamma + 1)
    return left, right


def max_depth(keys):
    """Internal nodes on the longest root-to-leaf path."""
    n = keys.shape[0]
    left, right = karras(keys)
    parent = np.full(2 * n - 1, -1, np.int64)
    parent[left] = np.arange(n - 1)
    parent[right] = np.arange(n - 1)
    assert (parent[1:] >= 0).all() and parent[0] == -1, "not one tree"
    depth = np.zeros(2 * n - 1, np.int64)
    anc = parent.copy()
    while (anc >= 0).any():
        m = anc >= 0
        depth[m] += 1
        anc[m] = parent[anc[m]]
    return int(depth[n - 1:].max())


def _stack_depth():
    with open(os.path.join(ROOT, "include", "rtspatial", "device", "tree.h")) as f:
        return int(re.search(r"kStackDepth\s*=\s*(\d+)", f.read()).group(1))


def test_split_rule_matches_a_balanced_tree_on_distinct_keys():
    keys = np.arange(1024, dtype=np.uint64)
    assert max_depth(keys) == 10


def test_deepest_trees_stay_within_the_bound():
    chain = np.array([0] + [1 << b for b in range(30)], np.uint64)
    for dups in (1, 2, 1000, 4096, 1 << 14):
        keys = np.sort(np.concatenate([chain, np.zeros(dups - 1, np.uint64)]))
        d = max_depth(keys)
        # the one-bit chain gives one level per key bit; the equal keys a
        # subtree of ceil(log2(dups)) levels under it
        assert d == 30 + int(np.ceil(np.log2(dups))), (dups, d)
        assert d <= 60
    # equal keys everywhere: only the positions split
    assert max_depth(np.zeros(1 << 14, np.uint64)) == 14
    # runs of equal keys at every level of the chain
    keys = np.sort(np.repeat(chain, 300))
    assert max_depth(keys) <= 60


def test_lane_stack_holds_the_deepest_walk():
    # The lane walk pops one entry and pushes at most two (both children),
    # so its stack never holds more than depth + 1 entries.
    assert _stack_depth() >= 60 + 2
