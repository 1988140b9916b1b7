"""Hand-built flat scenes where the order of a BVH traversal can decide the
closest hit (test helpers).

The reference traverses depth-first and prunes a box whose rounded entry
distance exceeds the closest hit found so far (src/renderer/transport.h:61-83).
When a primitive's computed hit distance lies in front of the rounded entry
distance of one of its boxes -- a triangle on a face of a tight box, hit where
it meets another at exactly the same distance -- which of the tied primitives
the reference returns depends on what it found before it reached that box.
These scenes make such cases common:

- grid_scene: a grid of unit squares in the plane z = 0, split into two
  triangles each, one triangle per leaf with a tight (flat) leaf box; branch
  boxes are the union of their children's, widened in z by `pad`.  edge_rays
  aims rays at the grid lines and diagonals, so that many pass through an
  edge two triangles share and hit both at the same distance.
- sphere_scene: spheres with tight leaf boxes (center +- radius, as the
  reference's own builder makes them) and grazing rays.
- fine_grid_scene: the same kind of grid at the scale of float ulps, seen by a
  camera whose rays land exactly on its edges, for renders.
- chain_scene: a maximally unbalanced hierarchy (every branch has a leaf and a
  subtree), for stack occupancy.

All scenes are valid flat scenes (src/flat/flat_validate.h).
"""
import numpy as np

import ptrace


def _material_arrays():
    m = np.zeros(1, ptrace.MATERIAL)
    m["type"] = ptrace.DIFFUSE
    m["texture_id"] = 0
    x = np.zeros(1, ptrace.TEXTURE)
    x["color"] = (0.5, 0.5, 0.5)
    return m, x


def _tree(boxes, prim_type, rng, pad=0.0, axes=(0, 1, 2), centers=None, swap=True):
    """Flat nodes over primitives with leaf boxes `boxes` (lo, hi): median
    splits on a random axis, children swapped at random; branch boxes are the
    union of their children's, widened by `pad` in z.  Returns (nodes, root,
    height)."""
    lo_all, hi_all = boxes
    cen = (lo_all + hi_all) / 2 if centers is None else centers
    nodes = []

    def build(ids):
        k = len(nodes)
        nodes.append(None)
        if len(ids) == 1:
            nodes[k] = (lo_all[ids[0]], hi_all[ids[0]], prim_type, -1, -1, int(ids[0]))
            return k, 1
        ax = axes[int(rng.integers(0, len(axes)))]
        srt = [ids[m] for m in np.argsort(cen[ids, ax], kind="stable")]
        h = len(srt) // 2
        left, hl = build(srt[:h])
        right, hr = build(srt[h:])
        if swap and rng.random() < 0.5:
            left, right = right, left
        lo = np.minimum(nodes[left][0], nodes[right][0]).astype(np.float32)
        hi = np.maximum(nodes[left][1], nodes[right][1]).astype(np.float32)
        lo[2] = np.float32(lo[2] - np.float32(pad))
        hi[2] = np.float32(hi[2] + np.float32(pad))
        nodes[k] = (lo, hi, ptrace.BRANCH, left, right, -1)
        return k, 1 + max(hl, hr)

    order = list(rng.permutation(len(lo_all)))
    root, height = build(order)
    out = np.zeros(len(nodes), ptrace.NODE)
    for k, (lo, hi, ty, left, right, prim) in enumerate(nodes):
        out[k]["minimum"] = lo
        out[k]["maximum"] = hi
        out[k]["type"] = ty
        out[k]["left"] = left
        out[k]["right"] = right
        out[k]["primitive"] = prim
    return out, root, height


def grid_scene(n=6, pad=1e-6, seed=0):
    rng = np.random.default_rng(seed)
    tris = []
    for i in range(n):
        for j in range(n):
            a, b, c, d = (i, j, 0), (i + 1, j, 0), (i + 1, j + 1, 0), (i, j + 1, 0)
            tris += [(a, b, c), (a, c, d)] if rng.random() < 0.5 else [(a, b, d), (b, c, d)]
    t = np.zeros(len(tris), ptrace.TRIANGLE)
    for k, (p, q, r) in enumerate(tris):
        t[k]["v0"], t[k]["v1"], t[k]["v2"] = p, q, r
        t[k]["uv1"], t[k]["uv2"] = (1, 0, 0), (0, 1, 0)
    v = np.stack([t["v0"], t["v1"], t["v2"]])
    nodes, root, height = _tree((v.min(0), v.max(0)), ptrace.TRIANGLE_PRIM, rng, pad=pad, axes=(0, 1),
                                centers=v.mean(0))
    m, x = _material_arrays()
    return ptrace.FlatScene(t, np.zeros(0, ptrace.SPHERE), nodes, m, x, np.zeros(0, np.uint8), root, height)


def fine_grid_scene(n=64, seed=0):
    """Camera rays that land on shared edges: a grid of n x n squares of side
    2^-20 (8 float ulps at 1.0) in the plane z = 1, two triangles each, with
    tight leaf boxes and branch boxes padded by one ulp (2^-23) in z; a 2-triangle area
    light above it; a pinhole camera looking straight down from 1 + 3 * 2^-16
    with a focus distance of 2^-16.  Every camera ray then has the same z
    direction, hits the plane at t = 3, and its x and y directions lie on the
    2^-23 lattice of numbers near 1, so a large share of the camera rays pass
    exactly through the grid lines and diagonals; there a box's entry distance
    (edge plane - origin) * (1/d) rounds to 3 or one ulp above it, so boxes
    whose entry distance exceeds the hit distance of their triangle are
    common."""
    s = 2.0 ** -20
    rng = np.random.default_rng(seed)
    tris = []
    for i in range(n):
        for j in range(n):
            a = (1 + i * s, 1 + j * s, 1.0)
            b = (1 + (i + 1) * s, 1 + j * s, 1.0)
            c = (1 + (i + 1) * s, 1 + (j + 1) * s, 1.0)
            d = (1 + i * s, 1 + (j + 1) * s, 1.0)
            tris += [(a, b, c), (a, c, d)] if rng.random() < 0.5 else [(a, b, d), (b, c, d)]
    focus = 2.0 ** -16
    height = 3 * focus
    lo, hi = 1 - 8 * s, 1 + (n + 8) * s
    light_z = 1 + 2 * height
    tris += [((lo, lo, light_z), (hi, hi, light_z), (hi, lo, light_z)),
             ((lo, lo, light_z), (lo, hi, light_z), (hi, hi, light_z))]
    t = np.zeros(len(tris), ptrace.TRIANGLE)
    for k, (p, q, r) in enumerate(tris):
        t[k]["v0"], t[k]["v1"], t[k]["v2"] = p, q, r
        t[k]["uv1"], t[k]["uv2"] = (1, 0, 0), (0, 1, 0)
    t["material_id"][-2:] = 1
    v = np.stack([t["v0"], t["v1"], t["v2"]])
    nodes, root, depth = _tree((v.min(0), v.max(0)), ptrace.TRIANGLE_PRIM, rng, pad=2.0 ** -23, axes=(0, 1, 2),
                               centers=v.mean(0))
    m = np.zeros(2, ptrace.MATERIAL)
    m["type"] = (ptrace.DIFFUSE, ptrace.LIGHT)
    m["texture_id"] = (0, 1)
    x = np.zeros(2, ptrace.TEXTURE)
    x["color"] = [(0.6, 0.5, 0.4), (15, 15, 15)]
    scene = ptrace.FlatScene(t, np.zeros(0, ptrace.SPHERE), nodes, m, x, np.zeros(0, np.uint8), root, depth)
    centre = 1 + n / 2 * s
    camera = ptrace.make_camera((centre, centre, 1 + height), (centre, centre, 1.0), (0, 1, 0), 60, 1.0, 0.0, focus)
    return scene, camera


def camera_rays(camera, width, height, count, seed):
    """Jittered camera rays with the renderer's arithmetic (transport.h:157-162,
    232-237 with a zero lens radius), in float32."""
    rng = np.random.default_rng(seed)
    f = np.float32
    x = rng.integers(0, width, count).astype(f)
    y = rng.integers(0, height, count).astype(f)
    u = ((x + rng.integers(0, 1 << 24, count).astype(f) * f(2.0 ** -24)) / f(width)).astype(f)
    v = ((y + rng.integers(0, 1 << 24, count).astype(f) * f(2.0 ** -24)) / f(height)).astype(f)
    c = {k: np.asarray(camera[k], f) for k in ("origin", "lower_left_corner", "horizontal", "vertical")}
    d = c["lower_left_corner"] + (u[:, None] * c["horizontal"]).astype(f)
    d = (d + (v[:, None] * c["vertical"]).astype(f)).astype(f)
    d = (d - c["origin"]).astype(f)
    rays = np.zeros(count, ptrace.RAY)
    rays["origin"] = c["origin"]
    rays["direction"] = d
    return rays


def edge_rays(n=6, count=20000, seed=0):
    """Rays from above the grid at points on its grid lines and diagonals."""
    rng = np.random.default_rng(seed)
    o = np.stack([rng.uniform(-1, n + 1, count), rng.uniform(-1, n + 1, count), rng.uniform(0.2, 5, count)], 1)
    kind = rng.integers(0, 3, count)
    u = rng.uniform(0, n, count)
    g = rng.integers(0, n + 1, count).astype(float)
    g2 = rng.integers(0, n, count).astype(float)
    f = rng.uniform(0, 1, count)
    tx = np.where(kind == 0, g, np.where(kind == 1, u, g2 + f))
    ty = np.where(kind == 0, u, np.where(kind == 1, g, g2 + (1 - f)))
    target = np.stack([tx, ty, np.zeros(count)], 1)
    rays = np.zeros(count, ptrace.RAY)
    rays["origin"] = o
    rays["direction"] = (target - o).astype(np.float32)
    return rays


def sphere_scene(count=40, seed=0):
    rng = np.random.default_rng(seed)
    s = np.zeros(count, ptrace.SPHERE)
    s["center"] = rng.uniform(-4, 4, (count, 3)).astype(np.float32)
    s["radius"] = rng.uniform(0.2, 1.0, count).astype(np.float32)
    c = s["center"]
    r = s["radius"][:, None]
    lo, hi = (c - r).astype(np.float32), (c + r).astype(np.float32)
    nodes, root, height = _tree((lo, hi), ptrace.SPHERE_PRIM, rng)
    m, x = _material_arrays()
    return ptrace.FlatScene(np.zeros(0, ptrace.TRIANGLE), s, nodes, m, x, np.zeros(0, np.uint8), root, height)


def grazing_rays(scene, count=20000, seed=0):
    """Rays from random origins aimed at points just inside or outside a
    sphere's silhouette."""
    rng = np.random.default_rng(seed)
    sp = scene.spheres[rng.integers(0, len(scene.spheres), count)]
    o = rng.uniform(-8, 8, (count, 3))
    to_c = sp["center"] - o
    side = np.cross(to_c, rng.normal(size=(count, 3)))
    side /= np.linalg.norm(side, axis=1, keepdims=True)
    reach = sp["radius"] * (1 + rng.uniform(-1e-5, 1e-5, count))
    target = sp["center"] + side * reach[:, None]
    rays = np.zeros(count, ptrace.RAY)
    rays["origin"] = o
    rays["direction"] = (target - o).astype(np.float32)
    return rays


def chain_scene(length=40, seed=0):
    """Triangles along a line; every branch holds one leaf and the rest."""
    rng = np.random.default_rng(seed)
    t = np.zeros(length, ptrace.TRIANGLE)
    x0 = np.arange(length, dtype=np.float32)
    t["v0"] = np.stack([x0, np.zeros(length), np.zeros(length)], 1)
    t["v1"] = np.stack([x0 + 1, np.zeros(length), np.zeros(length)], 1)
    t["v2"] = np.stack([x0, np.ones(length), rng.uniform(-0.5, 0.5, length)], 1)
    v = np.stack([t["v0"], t["v1"], t["v2"]])
    lo, hi = v.min(0) - np.float32(1e-3), v.max(0) + np.float32(1e-3)
    nodes = np.zeros(2 * length - 1, ptrace.NODE)
    # node 2k: branch over triangles k.., node 2k+1: leaf of triangle k; last node: leaf of the last triangle
    for k in range(length):
        leaf = 2 * k + 1 if k < length - 1 else 2 * k
        nodes[leaf]["type"] = ptrace.TRIANGLE_PRIM
        nodes[leaf]["primitive"] = k
        nodes[leaf]["minimum"], nodes[leaf]["maximum"] = lo[k], hi[k]
        nodes[leaf]["left"] = nodes[leaf]["right"] = -1
    for k in range(length - 2, -1, -1):
        b = 2 * k
        nodes[b]["type"] = ptrace.BRANCH
        leaf, rest = 2 * k + 1, 2 * k + 2
        nodes[b]["left"], nodes[b]["right"] = (leaf, rest) if rng.random() < 0.5 else (rest, leaf)
        nodes[b]["minimum"] = np.minimum(nodes[leaf]["minimum"], nodes[rest]["minimum"])
        nodes[b]["maximum"] = np.maximum(nodes[leaf]["maximum"], nodes[rest]["maximum"])
        nodes[b]["primitive"] = -1
    m, x = _material_arrays()
    return ptrace.FlatScene(t, np.zeros(0, ptrace.SPHERE), nodes, m, x, np.zeros(0, np.uint8), 0, length)


def random_rays(scene, count, seed):
    """Rays between random points of (a margin around) the scene's bounds."""
    rng = np.random.default_rng(seed)
    lo, hi = scene.nodes["minimum"].min(0), scene.nodes["maximum"].max(0)
    rays = np.zeros(count, ptrace.RAY)
    rays["origin"] = lo + (hi - lo) * rng.uniform(-0.2, 1.2, (count, 3))
    rays["direction"] = (lo + (hi - lo) * rng.uniform(0, 1, (count, 3)) - rays["origin"]).astype(np.float32)
    return rays


def oracle_hits(results):
    """(t bits, u bits, v bits, code) per ray of trace_result records, code =
    primitive*2 + sphere or -1: the CPU model's hit encoding."""
    r = np.asarray(results).view(ptrace.TRACE_RESULT)
    hit = r["hit"] == 1
    code = np.where(hit, r["rec"]["primitive_id"] * 2 + (r["rec"]["primitive_type"] == ptrace.SPHERE_PRIM), -1)
    out = np.zeros((len(r), 4), np.int32)
    out[:, 0] = np.where(hit, r["rec"]["t"].view(np.int32), 0)
    out[:, 1] = np.where(hit, r["rec"]["u"].view(np.int32), 0)
    out[:, 2] = np.where(hit, r["rec"]["v"].view(np.int32), 0)
    out[:, 3] = code
    return out
