"""Device linear BVH builder (pt_build_bvh) on the MI355X.

The reference builds its hierarchy on the host (src/bvh/bvh.h:132-171); the
device builder is an addition (a build in milliseconds), so it is held to the
flat format's own rules and to the renderer's exactness on its trees.  Bars:
  * the tree passes the oracle's restatement of flat_validate.h:10-83 with
    the depth pt_build_bvh reports;
  * every node equals tests/lbvh_model.py's (reference leaf boxes, Morton
    order, Karras splits, unions) byte for byte, up to the full 1,003,530
    triangle benchmark scene;
  * two builds of one input are byte-identical;
  * renders over device-built trees equal the oracle's renders over the same
    arrays (accumulation sums, PCG states, ray counts) bit for bit, and a
    batch over such trees equals single renderers;
  * depth 64 builds and depth 67 is refused (the 64-entry traversal stack,
    flat_data.h:52).
"""
import numpy as np
import pytest

import lbvh_model
import oracle
import ptrace

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def device():
    ok, name = ptrace.device_available()
    assert ok, f"no HIP device: {name}"
    print("device:", name, flush=True)


def _depth(nodes, root=0):
    stack, depth = [(root, 1)], 0
    while stack:
        i, d = stack.pop()
        depth = max(depth, d)
        if nodes[i]["type"] == ptrace.BRANCH:
            stack += [(int(nodes[i]["left"]), d + 1), (int(nodes[i]["right"]), d + 1)]
    return depth


def _check_build(triangles, spheres, materials=None, textures=None, pixels=None):
    scene, ms = ptrace.build_bvh_device(triangles, spheres, materials, textures, pixels)
    want = lbvh_model.build(scene.triangles, scene.spheres)
    assert scene.root == 0 and ms >= 0
    assert scene.nodes.tobytes() == want.tobytes()
    if materials is not None:
        assert oracle.validate(scene)
    return scene, ms


def _spheres(centers, radius=1.0):
    s = np.zeros(len(centers), ptrace.SPHERE)
    s["center"], s["radius"] = centers, radius
    return s


def test_room_tree_is_valid_and_renders_like_the_oracle():
    base, cam = ptrace.cornell(7, 96, 60)
    scene, _ = _check_build(base.triangles, base.spheres, base.materials, base.textures, base.pixels)
    assert scene.max_depth == _depth(scene.nodes)
    r = ptrace.Renderer(scene, cam, 96, 60, 5, False, 0, True, 8)
    r.add_samples(8)
    o = oracle.Renderer(scene, cam, 96, 60, 5, False, 0, True, 8)
    o.add_samples(8, threads=16)
    sums, rng = r.read_state()
    assert r.rays_traced() == o.rays_traced()
    assert np.array_equal(sums, o.sums()) and np.array_equal(rng, o.rng_state())


def test_benchmark_scene_tree_matches_the_model_and_renders_like_the_oracle():
    # The benchmark's icosphere (frequency 224: 1,003,520 mesh triangles,
    # 10 floor and light triangles), rebuilt on the device.
    base, cam = ptrace.icosphere(224, 7, 64, 36)
    scene, ms = _check_build(base.triangles, base.spheres, base.materials, base.textures, base.pixels)
    again, _ = ptrace.build_bvh_device(base.triangles, base.spheres, base.materials, base.textures, base.pixels)
    assert again.nodes.tobytes() == scene.nodes.tobytes()
    print(f"device build of {len(base.triangles)} triangles: {ms:.2f} ms, depth {scene.max_depth}", flush=True)
    assert ms > 0
    assert scene.max_depth == _depth(scene.nodes) <= 64
    r = ptrace.Renderer(scene, cam, 64, 36, 3, False, 0, True, 8)
    r.add_samples(2)
    o = oracle.Renderer(scene, cam, 64, 36, 3, False, 0, True, 8)
    o.add_samples(2, threads=16)
    sums, rng = r.read_state()
    assert r.rays_traced() == o.rays_traced()
    assert np.array_equal(sums, o.sums()) and np.array_equal(rng, o.rng_state())


@pytest.mark.parametrize("frequency", [8, 40])
def test_mesh_trees_match_the_model(frequency):
    base, cam = ptrace.icosphere(frequency, 11, 32, 18)
    scene, _ = _check_build(base.triangles, base.spheres, base.materials, base.textures, base.pixels)
    rays = _camera_rays(cam, 32, 18)
    got, boxes, prims = ptrace.query_rays(scene, rays, stats=True)
    want, want_boxes, want_prims = oracle.query(scene, rays)
    assert got.tobytes() == want.tobytes()
    assert np.array_equal(boxes, want_boxes) and np.array_equal(prims, want_prims)
    assert ptrace.query_rays_wide(scene, rays)[0].tobytes() == want.tobytes()


def _camera_rays(cam, w, h):
    c = np.asarray(cam).reshape(())
    s = (np.arange(w, dtype=np.float32) + np.float32(0.5)) / np.float32(w)
    t = (np.arange(h, dtype=np.float32) + np.float32(0.5)) / np.float32(h)
    ss, tt = np.meshgrid(s, t)
    rays = np.zeros(w * h, ptrace.RAY)
    rays["origin"] = c["origin"]
    target = (c["lower_left_corner"] + ss.reshape(-1, 1) * c["horizontal"] + tt.reshape(-1, 1) * c["vertical"])
    rays["direction"] = (target - c["origin"]).astype(np.float32)
    return rays


def test_small_and_degenerate_inputs():
    tri = ptrace.cornell(7, 8, 8)[0].triangles
    for t, s in [(tri[:1], None), (None, _spheres([(1, 2, 3)])), (tri[:1], _spheres([(0, 0, 0)])), (tri[:2], None)]:
        scene, _ = _check_build(t, s)
        assert scene.max_depth == _depth(scene.nodes)
    # Equal codes: 4096 copies of one triangle split by position, depth 13.
    scene, _ = _check_build(np.repeat(tri[:1], 4096), None)
    assert scene.max_depth == 13
    empty, ms = ptrace.build_bvh_device(None, None)
    assert empty.root == -1 and empty.max_depth == 0 and len(empty.nodes) == 0 and ms == 0


def _chain(origins):
    # Sphere centres whose Morton codes are 0 and every power of two: each
    # split of the radix tree peels one leaf, a chain of 63 levels above the
    # origin's leaves.
    centres = [(0, 0, 0)] * origins + [(2 ** b, 0, 0) for b in range(21)] + \
              [(0, 2 ** b, 0) for b in range(21)] + [(0, 0, 2 ** b) for b in range(21)] + [(2 ** 21,) * 3]
    return _spheres(centres)


def test_depth_limit():
    scene, _ = _check_build(None, _chain(1))
    assert scene.max_depth == 64
    with pytest.raises(RuntimeError, match="deeper than 64"):
        ptrace.build_bvh_device(None, _chain(8))


def test_batch_over_device_built_trees_matches_single_renderers():
    # Device-built hierarchies in one batched launch: every view equals a
    # single renderer over the same tree (and so the oracle, by the tests
    # above), and the batch's ray count their sum.
    built = [ptrace.cornell(3, 40, 30), ptrace.soup(6000, 2, 48, 32), ptrace.icosphere(20, 5, 36, 36)]
    scenes = [_check_build(s.triangles, s.spheres, s.materials, s.textures, s.pixels)[0] for s, _ in built]
    views = [(0, built[0][1], 40, 30, 7), (1, built[1][1], 48, 32, 9), (2, built[2][1], 36, 36, 3)]
    batch = ptrace.Batch(scenes, views, direct_light=True, max_depth=6)
    batch.add_samples(3)
    total = 0
    for k, (si, cam, w, h, seed) in enumerate(views):
        r = ptrace.Renderer(scenes[si], cam, w, h, seed, False, ptrace.DEBUG_SHADED, True, 6)
        r.add_samples(3)
        assert np.array_equal(batch.image(k), r.image()), k
        total += r.rays_traced()
    assert batch.stats()["rays"] == total
