"""Numpy model of the device linear BVH builder (pt_build_bvh, pt_lbvh.hip).

Test infrastructure: the GPU build is compared with this model node for
node.  The reference has no linear builder (its hierarchy is built on the
host, src/bvh/bvh.h:132-171), so the model restates the published method
(Karras, "Maximizing Parallelism in the Construction of BVHs, Octrees, and
k-d Trees", HPG 2012) over the reference's leaf boxes:
  * leaf boxes: triangle.h:96-109 (per-axis min/max, minus/plus 1e-4, one
    float step outwards), sphere.h:69-75 (center -/+ radius), all in float32;
  * centroids 0.5*lo + 0.5*hi, 21 bits per axis of their position in the
    centroids' bounds, interleaved x-y-z from the top bit into a 63-bit code,
    a stable sort of the codes;
  * internal node i covers the run of sorted leaves that share a longer
    common prefix with leaf i than its neighbour on the other side does, and
    splits it where the prefix shared with leaf i ends (equal codes extended
    by the leaf positions);
  * branch boxes are ffmin/ffmax unions of the children (aabb.h:40-56).
Pinned by `top_down_ranges` (the binary radix tree of the same keys built by
recursive splitting at the highest differing bit, which Karras's rule must
reproduce) and by the oracle's flat-scene validator in the tests.
"""
import numpy as np

import ptrace

F32 = np.float32


def leaf_boxes(triangles, spheres):
    """(lo, hi) float32 arrays of shape (n, 3): triangles first, then spheres."""
    t = np.asarray(triangles)
    s = np.asarray(spheres)
    v0, v1, v2 = (np.asarray(t[k], F32).reshape(-1, 3) for k in ("v0", "v1", "v2"))
    # ffmin(v0, ffmin(v1, v2)) and ffmax alike; for equal operands either one
    # gives the same box after the 1e-4 offset.
    mn = np.where(v0 < np.where(v1 < v2, v1, v2), v0, np.where(v1 < v2, v1, v2))
    mx = np.where(v0 > np.where(v1 > v2, v1, v2), v0, np.where(v1 > v2, v1, v2))
    tlo = np.nextafter((mn - F32(0.0001)).astype(F32), F32(-np.inf))
    thi = np.nextafter((mx + F32(0.0001)).astype(F32), F32(np.inf))
    c = np.asarray(s["center"], F32).reshape(-1, 3)
    r = np.asarray(s["radius"], F32).reshape(-1, 1)
    return (np.concatenate([tlo, (c - r).astype(F32)]).astype(F32),
            np.concatenate([thi, (c + r).astype(F32)]).astype(F32))


def _spread21(x):
    x = x.astype(np.uint64) & np.uint64(0x1FFFFF)
    out = np.zeros_like(x)
    for b in range(21):
        out |= ((x >> np.uint64(b)) & np.uint64(1)) << np.uint64(3 * b)
    return out


def morton_codes(lo, hi):
    centroid = (F32(0.5) * lo + F32(0.5) * hi).astype(F32)
    cmin, cmax = centroid.min(axis=0), centroid.max(axis=0)
    ext = (cmax - cmin).astype(F32)
    with np.errstate(divide="ignore", invalid="ignore"):
        x = np.where(ext > 0, ((centroid - cmin).astype(F32) / ext).astype(F32), F32(0))
    x = np.clip(x, F32(0), F32(1))
    q = np.minimum((x * F32(2097152.0)).astype(F32), F32(2097151.0)).astype(np.uint32)
    return (_spread21(q[:, 0]) << np.uint64(2)) | (_spread21(q[:, 1]) << np.uint64(1)) | _spread21(q[:, 2])


def _bit_length(x):
    x = x.astype(np.uint64)
    high = (x >> np.uint64(32)).astype(np.float64)
    low = (x & np.uint64(0xFFFFFFFF)).astype(np.float64)
    return np.where(high > 0, 32 + np.frexp(high)[1], np.frexp(low)[1]).astype(np.int64)


def _delta(keys, i, j):
    """Common prefix length of sorted leaves i and j (64 + that of the
    positions where the codes are equal); -1 outside the leaves."""
    n = len(keys)
    inside = (j >= 0) & (j < n)
    jc = np.clip(j, 0, n - 1)
    a, b = keys[i], keys[jc]
    same = a == b
    d = np.where(same, 64 + 64 - _bit_length(i.astype(np.uint64) ^ jc.astype(np.uint64)), 64 - _bit_length(a ^ b))
    return np.where(inside, d, -1)


def internal_nodes(keys):
    """(left, right, first, last, split) of the n - 1 internal nodes; children in
    flat numbering (internal k -> k, sorted leaf k -> n - 1 + k)."""
    n = len(keys)
    i = np.arange(n - 1, dtype=np.int64)
    d = np.where(_delta(keys, i, i + 1) > _delta(keys, i, i - 1), 1, -1)
    dmin = _delta(keys, i, i - d)
    lmax = np.full(n - 1, 2, np.int64)
    grow = _delta(keys, i, i + lmax * d) > dmin
    while grow.any():
        lmax = np.where(grow, lmax * 2, lmax)
        grow = grow & (_delta(keys, i, i + lmax * d) > dmin)
    length = np.zeros(n - 1, np.int64)
    t = lmax >> 1
    while (t >= 1).any():
        step = (t >= 1) & (_delta(keys, i, i + (length + t) * d) > dmin)
        length = np.where(step, length + t, length)
        t = t >> 1
    j = i + length * d
    dnode = _delta(keys, i, j)
    s = np.zeros(n - 1, np.int64)
    t = length.copy()
    live = np.ones(n - 1, bool)
    while live.any():
        t = np.where(live, (t + 1) >> 1, t)
        step = live & (_delta(keys, i, i + (s + t) * d) > dnode)
        s = np.where(step, s + t, s)
        live = live & (t > 1)
    split = i + s * d + np.where(d < 0, -1, 0)
    first, last = np.minimum(i, j), np.maximum(i, j)
    left = np.where(first == split, n - 1 + split, split)
    right = np.where(last == split + 1, n - 1 + split + 1, split + 1)
    return left, right, first, last, split


def top_down_ranges(keys):
    """{(first, last, split)} of the binary radix tree of the keys by
    recursive splitting (equal codes extended by position)."""
    ext = [(int(k) << 32) | p for p, k in enumerate(keys)]
    out = set()
    stack = [(0, len(keys) - 1)]
    while stack:
        a, b = stack.pop()
        if a == b:
            continue
        bit = (ext[a] ^ ext[b]).bit_length() - 1
        split = next(k for k in range(a, b) if (ext[k + 1] >> bit) & 1)
        out.add((a, b, split))
        stack += [(a, split), (split + 1, b)]
    return out


def build(triangles, spheres):
    """The flat nodes (ptrace.NODE) of the device builder, root 0."""
    lo, hi = leaf_boxes(triangles, spheres)
    n, nt = len(lo), len(np.asarray(triangles))
    codes = morton_codes(lo, hi)
    order = np.argsort(codes, kind="stable")
    keys = codes[order]
    nodes = np.zeros(2 * n - 1, ptrace.NODE)
    leaves = nodes[n - 1:]
    leaves["minimum"], leaves["maximum"] = lo[order], hi[order]
    leaves["type"] = np.where(order < nt, ptrace.TRIANGLE_PRIM, ptrace.SPHERE_PRIM)
    leaves["left"] = leaves["right"] = -1
    leaves["primitive"] = np.where(order < nt, order, order - nt)
    if n == 1:
        return nodes
    left, right = internal_nodes(keys)[:2]
    inner = nodes[: n - 1]
    inner["type"], inner["left"], inner["right"], inner["primitive"] = ptrace.BRANCH, left, right, -1
    # Boxes bottom up, deepest level of internal nodes first.
    frontier = np.array([0])
    levels = []
    while len(frontier):
        levels.append(frontier)
        frontier = np.concatenate([left[frontier], right[frontier]])
        frontier = frontier[frontier < n - 1]
    for level in reversed(levels):
        a, b = nodes[left[level]], nodes[right[level]]
        nodes["minimum"][level] = np.where(a["minimum"] < b["minimum"], a["minimum"], b["minimum"])
        nodes["maximum"][level] = np.where(a["maximum"] > b["maximum"], a["maximum"], b["maximum"])
    return nodes
