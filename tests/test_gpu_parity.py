"""HIP core against the oracle and the reference's fixtures (needs an MI355X).

Bars (written per test):
  * closest-hit records, traversal counts and the diagnostic views: bit-exact;
  * sphere texture coordinates in hit records: within 1.2e-7 absolute (the
    reference's host build calls glibc atan2f/asinf, the device evaluates
    atan2/asin in double and rounds; one float ulp of an angle below pi,
    divided by 2*pi or pi).  They feed image textures on spheres only, which
    none of the shaded scenes below has;
  * shaded images, accumulation sums, PCG states, RGBA bytes and ray counts:
    bit-exact on every pixel.
"""
import numpy as np
import pytest

import oracle
import ptrace
from conftest import golden, result_bytes

pytestmark = pytest.mark.gpu

ASSETS = ["asset_cube.npz", "asset_triangle.npz", "asset_sphere.npz", "cornell.npz", "pruning.npz"]


@pytest.fixture(scope="module", autouse=True)
def device():
    ok, name = ptrace.device_available()
    assert ok, f"no HIP device: {name}"
    print("device:", name, flush=True)
    return name


def _ulp_diff(a, b):
    a = np.asarray(a, np.float32).view(np.int32).astype(np.int64)
    b = np.asarray(b, np.float32).view(np.int32).astype(np.int64)
    return np.abs(a - b)


def _compare_results(got, want):
    """Bit-exact except sphere u,v (absolute 1.2e-7)."""
    got, want = got.view(ptrace.TRACE_RESULT), want.view(ptrace.TRACE_RESULT)
    assert np.array_equal(got["hit"], want["hit"]) and np.array_equal(got["error"], want["error"])
    sphere = want["rec"]["primitive_type"] == ptrace.SPHERE_PRIM
    for f in ("t", "p", "normal", "material_id", "primitive_type", "primitive_id"):
        assert np.array_equal(got["rec"][f], want["rec"][f]), f
    for f in ("u", "v"):
        assert np.array_equal(got["rec"][f][~sphere], want["rec"][f][~sphere]), f
        diff = np.abs(got["rec"][f][sphere].astype(np.float64) - want["rec"][f][sphere])
        assert diff.max(initial=0) <= 1.2e-7, f


@pytest.mark.parametrize("name", ASSETS)
def test_query_matches_reference(name):
    scene, _, g = golden(name)
    for brute in (0, 1):
        key = f"query_b{brute}"
        if key not in g:
            continue
        got, boxes, prims = ptrace.query_rays(scene, g["rays"], bool(brute), stats=True)
        _compare_results(got, g[key])
        assert np.array_equal(boxes, g[f"boxes_b{brute}"]) and np.array_equal(prims, g[f"prims_b{brute}"])


def test_query_pruning_tie_rule():
    scene, _, g = golden("pruning.npz")
    got, _, prims = ptrace.query_rays(scene, g["rays"], False, stats=True)
    assert got["hit"][0] == 1 and got["rec"]["material_id"][0] == 1 and prims[0] == 2


def _one_triangle_scene(v0, v1, v2, normals=None):
    t = np.zeros(1, ptrace.TRIANGLE)
    t["v0"], t["v1"], t["v2"] = v0, v1, v2
    t["uv0"], t["uv1"], t["uv2"] = (0, 0, 0), (1, 0, 0), (0, 1, 0)
    if normals is not None:
        t["n0"], t["n1"], t["n2"] = normals
        t["smooth"] = 1
    tex = np.zeros(1, ptrace.TEXTURE)
    tex["color"], tex["scale"], tex["clamp"], tex["image_id"] = (0.5, 0.5, 0.5), (1, 1, 0), 1, -1
    mat = np.zeros(1, ptrace.MATERIAL)
    mat["texture_id"] = 0
    return ptrace.build_scene(t, None, mat, tex)


def test_query_analytic_triangle_cases_and_oracle():
    # tests/triangle_tests.cpp:40-87 through the device query path; the
    # query interval is fixed at [0.001, FLT_MAX) (cuda_renderer.cu:117).
    scene = _one_triangle_scene((0, 0, 0), (1, 0, 0), (0, 1, 0))
    cases = [((0.25, 0.25, 2), (0, 0, -2), True), ((0.25, 0.25, -2), (0, 0, 2), True),
             ((0.8, 0.8, 1), (0, 0, -1), False), ((-0.1, 0.2, 1), (0, 0, -1), False),
             ((0.2, 0.2, 1), (1, 0, 0), False), ((0.2, 0.2, 0), (1, 0, 0), False),
             ((0.2, 0.2, 1), (0, 0, 0), False), ((0.2, 0.2, 1), (0, 0, 1), False),
             ((0, 0, 1), (0, 0, -1), True), ((1, 0, 1), (0, 0, -1), True), ((0, 1, 1), (0, 0, -1), True),
             ((0.5, 0.5, 1), (0, 0, -1), True), ((0.5, 0, 1), (0, 0, -1), True),
             ((0.25, 0.25, 1), (0, 0, -1e-8), True)]
    rays = np.zeros(len(cases), ptrace.RAY)
    for i, (o, d, _) in enumerate(cases):
        rays[i]["origin"], rays[i]["direction"] = o, d
    got = ptrace.query_rays(scene, rays)
    assert [bool(h) for h in got["hit"]] == [c[2] for c in cases]
    assert got["rec"]["t"][0] == 1 and tuple(got["rec"]["p"][0]) == (0.25, 0.25, 0)
    assert tuple(got["rec"]["normal"][0]) == (0, 0, 1) and tuple(got["rec"]["normal"][1]) == (0, 0, 1)
    want, _, _ = oracle.query(scene, rays)
    assert np.array_equal(result_bytes(got), result_bytes(want))


def test_query_random_scenes_match_oracle():
    rng = np.random.default_rng(3)
    for n, seed in ((1, 1), (7, 2), (500, 3), (20000, 4)):
        scene, cam = ptrace.soup(n, seed, 64, 48)
        rays = np.zeros(4096, ptrace.RAY)
        rays["origin"] = rng.uniform(-3, 3, (4096, 3)).astype(np.float32)
        rays["direction"] = rng.normal(size=(4096, 3)).astype(np.float32)
        rays["direction"][::5, :2] = 0
        for brute in ((0, 1) if n <= 500 else (0,)):
            got, gb, gp = ptrace.query_rays(scene, rays, bool(brute), stats=True)
            want, wb, wp = oracle.query(scene, rays, bool(brute))
            assert np.array_equal(result_bytes(got), result_bytes(want)), (n, brute)
            assert np.array_equal(gb, wb) and np.array_equal(gp, wp)


def test_query_empty_inputs():
    scene, _, g = golden("asset_cube.npz")
    assert len(ptrace.query_rays(scene, np.zeros(0, ptrace.RAY))) == 0
    empty = ptrace.FlatScene(materials=scene.materials, textures=scene.textures)
    got = ptrace.query_rays(empty, g["rays"][:10])
    assert not got["hit"].any() and (got["rec"]["material_id"] == -1).all()


def _render(scene, cam, w, h, seed, batches, mode, dl, brute=False, max_depth=0, block=None):
    r = ptrace.Renderer(scene, cam, w, h, seed, brute, mode, dl, max_depth)
    if block:
        r.set_block_shape(*block)
    for b in batches:
        r.add_samples(b)
    return r


def _image_close(got, want):
    same = (got == want).all(axis=1)
    rel = np.sqrt(((got.astype(np.float64) - want) ** 2).sum() / max(1e-20, (want.astype(np.float64) ** 2).sum()))
    return same.mean(), rel


RENDER = [("asset_cube.npz", 19, 13, 7, [5, 11]), ("asset_sphere.npz", 19, 13, 7, [5, 11]),
          ("asset_triangle.npz", 19, 13, 7, [5, 11]), ("cornell.npz", 40, 25, 7, [3, 5])]


@pytest.mark.parametrize("name,w,h,seed,batches", RENDER)
def test_render_matches_reference_fixture(name, w, h, seed, batches):
    scene, cam, g = golden(name)
    cam = g.get("lens_camera", cam)
    for key in sorted(k for k in g if k.startswith("image_m")):
        mode, dl = int(key[7]), bool(int(key[-1]))
        r = _render(scene, cam, w, h, seed, batches, mode, dl)
        assert r.rays_traced() == int(g[key.replace("image", "rays")]), key
        assert np.array_equal(r.image(), g[key]), key


def test_sphere_lens_scene_and_brute_mode():
    scene, cam, g = golden("spheres.npz")
    r = _render(scene, cam, 31, 21, 19, [16, 48], 0, False)
    assert np.array_equal(r.image(), g["image_m0_d0"])
    b = _render(scene, cam, 31, 21, 19, [64], 0, False, brute=True)
    assert np.array_equal(b.image(), r.image()) and b.rays_traced() == r.rays_traced()
    for mode in range(1, 6):
        d = _render(scene, cam, 31, 21, 19, [64], mode, False)
        assert np.array_equal(d.image(), g[f"image_m{mode}_d0"]), mode


@pytest.mark.parametrize("occluded", [0, 1])
def test_triangle_light_mis(occluded):
    scene, cam, g = golden(f"triangle_light_{occluded}.npz")
    r = _render(scene, cam, 19, 13, 11, [16], 0, True)
    assert np.array_equal(r.image(), g["image_m0_d1"]) and r.rays_traced() == int(g["rays_m0_d1"])


@pytest.mark.parametrize("mode", range(6))
def test_batching_reset_and_block_shapes(mode):
    scene, cam, _ = golden("asset_sphere.npz")
    a = _render(scene, cam, 19, 13, 7, [5, 11], mode, True, block=(8, 8))
    b = _render(scene, cam, 19, 13, 7, [16], mode, True, block=(32, 8))
    assert np.array_equal(a.image(), b.image()) and a.rays_traced() == b.rays_traced()
    sa, ra = a.read_state()
    sb, rb = b.read_state()
    assert np.array_equal(sa, sb) and np.array_equal(ra, rb)
    a.reset(cam)
    assert a.sample_count() == 0 and a.rays_traced() == 0 and not a.image().any()
    a.add_samples(16)
    assert np.array_equal(a.image(), b.image())


def test_config1_room_matches_oracle_state():
    # Cornell room, 320x200, 16 spp, path depth 8, MIS: accumulation sums and
    # PCG states against the oracle for every pixel.
    scene, cam = ptrace.cornell(7, 320, 200)
    r = _render(scene, cam, 320, 200, 7, [16], 0, True, max_depth=8)
    o = oracle.Renderer(scene, cam, 320, 200, 7, False, 0, True, 8)
    o.add_samples(16, threads=16)
    gs, grng = r.read_state()
    assert r.rays_traced() == o.rays_traced()
    assert np.array_equal(gs, o.sums())
    assert np.array_equal(grng, o.rng_state())
    assert np.array_equal(r.image(), o.image())
    assert np.array_equal(r.rgba_image(), o.rgba_image())


def test_direct_lighting_only_per_pixel():
    # max_depth = 1: NEE at the first hit plus the BSDF-sampled light hit.
    # The direct-lighting bar is 1e-5 relative per pixel (and exactly 0
    # where the oracle is 0); the renderer meets it with equality.
    scene, cam = ptrace.cornell(11, 160, 100)
    r = _render(scene, cam, 160, 100, 3, [8], 0, True, max_depth=1)
    o = oracle.Renderer(scene, cam, 160, 100, 3, False, 0, True, 1)
    o.add_samples(8, threads=16)
    got, want = r.image().astype(np.float64), o.image().astype(np.float64)
    rel = np.abs(got - want) / np.maximum(np.abs(want), 1e-30)
    assert (rel[want != 0] <= 1e-5).all()
    assert (got[want == 0] == 0).all()
    assert np.array_equal(r.image(), o.image())
    assert r.rays_traced() == o.rays_traced()


def test_rgba_and_image_division_match_oracle():
    scene, cam, _ = golden("cornell.npz")
    r = _render(scene, cam, 40, 25, 7, [3], 0, True)
    o = oracle.Renderer(scene, cam, 40, 25, 7, False, 0, True, 50)
    o.add_samples(3)
    assert np.array_equal(r.read_state()[0], o.sums())
    assert np.array_equal(r.image(), o.image())
    assert np.array_equal(r.rgba_image(), o.rgba_image())


def test_error_paths_and_ownership():
    scene, cam, _ = golden("asset_cube.npz")
    r = ptrace.Renderer(scene, cam, 7, 5, 1)
    with pytest.raises(ValueError):
        r.add_samples(0)
    with pytest.raises(ValueError):
        r.set_block_shape(3, 3)
    bad_cam = np.array(cam, copy=True)
    bad_cam["lens_radius"] = -1
    with pytest.raises(ValueError):
        r.reset(bad_cam)
    del scene
    r.add_samples(2)
    assert r.image().shape == (35, 3) and r.sample_count() == 2
    with pytest.raises(ValueError):
        ptrace.Renderer(golden("asset_cube.npz")[0], cam, 0, 5, 1)
    broken = golden("asset_cube.npz")[0]
    broken.nodes["left"][0] = 999
    with pytest.raises(RuntimeError):
        ptrace.Renderer(broken, cam, 7, 5, 1)


def test_state_roundtrip_resumes_exactly():
    scene, cam, _ = golden("cornell.npz")
    a = _render(scene, cam, 40, 25, 7, [6], 0, True)
    b = _render(scene, cam, 40, 25, 7, [2], 0, True)
    s, rng = b.read_state()
    c = ptrace.Renderer(scene, cam, 40, 25, 7, False, 0, True)
    c.write_state(s, rng, b.sample_count(), b.rays_traced())
    c.add_samples(4)
    assert np.array_equal(a.image(), c.image()) and a.rays_traced() == c.rays_traced()


def test_wide_traversal_matches_reference_order_on_benchmark_scene():
    # The same comparison on the benchmark's 1,003,530-triangle scene (its
    # camera at 960x540, 8 spp): about 11M queries, compared through the
    # accumulation sums, the PCG states and the ray count.
    scene, cam = ptrace.icosphere(224, 7, 960, 540)
    states = []
    for exact in (True, False):
        r = ptrace.Renderer(scene, cam, 960, 540, 7, False, 0, True, 8)
        r.set_tuning(exact_traversal=exact)
        r.add_samples(8)
        states.append((r.read_state(), r.rays_traced()))
    (s0, g0), rays0 = states[0]
    (s1, g1), rays1 = states[1]
    assert rays0 == rays1 and rays0 > 10_000_000
    assert np.array_equal(g0, g1)
    assert np.array_equal(s0.view(np.uint32), s1.view(np.uint32))


@pytest.mark.parametrize("maker", ["cornell", "icosphere", "soup", "asset_sphere", "icosphere_dx0"])
def test_wide_traversal_matches_reference_order(maker):
    # The shaded path's 4-wide nearest-first traversal must return the same
    # closest hit as the reference-order binary traversal on every query:
    # accumulation sums, PCG states and ray counts are compared bit for bit.
    # "icosphere_dx0": every camera ray has a direction x component of exactly
    # 0 (the camera's corner, horizontal and vertical vectors and its origin
    # share x = 0), so the box test sees 1/d = +-inf on that axis for every
    # primary ray (the secondary rays have no zero component).
    if maker == "cornell":
        scene, cam = ptrace.cornell(5, 96, 64)
    elif maker == "icosphere":
        scene, cam = ptrace.icosphere(40, 3, 96, 64)
    elif maker == "icosphere_dx0":
        scene, cam = ptrace.icosphere(40, 3, 96, 64)
        cam = np.array(cam, copy=True)
        cam["origin"] = (0.0, 0.55, 3.1)
        cam["lower_left_corner"] = (0.0, -0.55, 1.85)
        cam["horizontal"] = (0.0, 0.0, 0.5)
        cam["vertical"] = (0.0, 2.2, 0.0)
    elif maker == "soup":
        scene, cam = ptrace.soup(30000, 2, 96, 64)
    else:
        scene, cam, _ = golden("asset_sphere.npz")
    states = []
    for exact in (True, False):
        r = ptrace.Renderer(scene, cam, 96, 64, 9, False, 0, True, 8)
        r.set_tuning(exact_traversal=exact)
        r.add_samples(3)
        r.add_samples(5)
        states.append((r.read_state(), r.rays_traced()))
    (s0, g0), rays0 = states[0]
    (s1, g1), rays1 = states[1]
    assert rays0 == rays1
    assert np.array_equal(s0, s1) and np.array_equal(g0, g1)


@pytest.mark.parametrize("threshold", [-1, 1000, 1 << 30])
def test_tail_drain_matches_wavefront_loop(threshold):
    # The tail kernel (each remaining pixel finished in one thread) must give
    # the same sums, PCG states and ray counts as running the bounce loop to
    # the end; -1 never drains, 2^30 drains at the first check.
    scene, cam = ptrace.icosphere(30, 4, 128, 72)
    ref = ptrace.Renderer(scene, cam, 128, 72, 5, False, 0, True, 8)
    ref.set_tuning(drain_threshold=-1)
    ref.add_samples(6)
    r = ptrace.Renderer(scene, cam, 128, 72, 5, False, 0, True, 8)
    r.set_tuning(drain_threshold=threshold, check_interval=2)
    r.add_samples(2)
    r.add_samples(4)
    (a, ga), (b, gb) = ref.read_state(), r.read_state()
    assert np.array_equal(a, b) and np.array_equal(ga, gb)
    assert ref.rays_traced() == r.rays_traced()
    if threshold > 0:
        assert r.last_launch_stats()["drained_paths"] > 0


def test_partitions_split_path_lanes_match_one_pass():
    # Two pixel partitions of about 92,000 pixels each on one 256-lane block per
    # compute unit: each partition's path-lane launch of 6 samples runs in two
    # passes of its own (4 samples, then the rest in cost order, with its own
    # cost, key and queue buffers); the result must equal one partition in one
    # pass.
    scene, cam = ptrace.icosphere(30, 6, 460, 400)
    ref = ptrace.Renderer(scene, cam, 460, 400, 13, False, 0, True, 8)
    ref.set_tuning(partitions=1, path_first_samples=-1)
    ref.add_samples(2)
    ref.add_samples(6)
    r = ptrace.Renderer(scene, cam, 460, 400, 13, False, 0, True, 8)
    r.set_tuning(partitions=2, path_blocks_per_cu=1)
    r.add_samples(2)
    r.add_samples(6)
    assert r.last_launch_stats()["split_launches"] == 2
    (a, ga), (b, gb) = ref.read_state(), r.read_state()
    assert np.array_equal(a, b) and np.array_equal(ga, gb)
    assert ref.rays_traced() == r.rays_traced()


@pytest.mark.parametrize("parts", [2, 3, 5])
def test_pixel_partitions_match_single_loop(parts):
    # Concurrent pixel partitions (4096-pixel chunks dealt round-robin, the
    # last chunk partial) must reproduce the single bounce loop exactly, also
    # when the partition count changes between add_samples calls.
    scene, cam = ptrace.icosphere(30, 6, 200, 150)
    ref = ptrace.Renderer(scene, cam, 200, 150, 11, False, 0, True, 8)
    ref.set_tuning(partitions=1)
    ref.add_samples(3)
    ref.add_samples(2)
    r = ptrace.Renderer(scene, cam, 200, 150, 11, False, 0, True, 8)
    r.set_tuning(partitions=parts, drain_threshold=2000)
    r.add_samples(3)
    ext, sh = r.iteration_log()
    st = r.last_launch_stats()
    assert int(ext.sum()) + int(sh.sum()) <= st["extend_rays"] + st["shadow_rays"]
    r.set_tuning(partitions=1 if parts != 1 else 2)
    r.add_samples(2)
    (a, ga), (b, gb) = ref.read_state(), r.read_state()
    assert np.array_equal(a, b) and np.array_equal(ga, gb)
    assert ref.rays_traced() == r.rays_traced()
    r.reset(cam)
    r.set_tuning(partitions=parts)
    r.add_samples(5)
    assert np.array_equal(ref.image(), r.image()) and ref.rays_traced() == r.rays_traced()


@pytest.mark.parametrize("steps,drain", [(1, 500), (4, 20000), (2, -1)])
def test_handoff_leaves_results_unchanged(steps, drain):
    # Queries still running a few steps after a trace launch's queues are
    # empty are saved and resumed by the next launch, their paths shaded one
    # iteration later.  Pixels, PCG states and ray counts must not change,
    # with the tail kernel taking over while queries are saved (drain 20000)
    # or never (-1).
    scene, cam = ptrace.icosphere(40, 3, 256, 192)
    ref = ptrace.Renderer(scene, cam, 256, 192, 17, False, 0, True, 8)
    ref.set_tuning(handoff_steps=-1, drain_threshold=drain)
    ref.add_samples(4)
    r = ptrace.Renderer(scene, cam, 256, 192, 17, False, 0, True, 8)
    r.set_tuning(handoff_steps=steps, drain_threshold=drain)
    r.add_samples(4)
    (a, ga), (b, gb) = ref.read_state(), r.read_state()
    assert np.array_equal(a, b) and np.array_equal(ga, gb)
    assert ref.rays_traced() == r.rays_traced()
    st, sr = r.last_launch_stats(), ref.last_launch_stats()
    assert (st["extend_rays"], st["shadow_rays"]) == (sr["extend_rays"], sr["shadow_rays"])
    ext, sh = r.iteration_log()
    assert int(ext.sum()) + int(sh.sum()) <= st["extend_rays"] + st["shadow_rays"]


@pytest.mark.parametrize("scene_name,blocks,batch,first,group,steps,split", [("icosphere", 1, 1, 2, 64, 1, True),
                                                                              ("icosphere", -1, 16, 4, 64, 8, False),
                                                                              ("cornell", 1, 64, 1, 1, 3, True),
                                                                              ("cornell", 2, 16, -1, 64, 1, False),
                                                                              ("icosphere", 1, 24, 4, 7, 8, True),
                                                                              ("cornell", 1, 24, 4, 4096, 64, True)])
def test_path_lanes_match_wavefront(scene_name, blocks, batch, first, group, steps, split):
    # The path-lane kernel (every remaining pixel on a lane's state machine,
    # shading batched per wave, lanes taking pixel after pixel) must equal the
    # wavefront bounce loop exactly: pixels, PCG states and ray counts.  With
    # more samples in a call than `first` and more pixels than lanes in flight
    # (76800 pixels; one 256-lane block per compute unit), the launch runs in
    # two passes: `first` samples of every pixel, then the rest with groups of
    # `group` pixels reordered by cost (7 and 4096 leave a partial last
    # group).  With one lane per pixel or 2 blocks per compute unit every
    # pixel has a lane from the start: one pass.  `steps`: traversal steps a
    # tracing lane takes per pass of its state machine.
    if scene_name == "icosphere":
        scene, cam = ptrace.icosphere(40, 3, 320, 240)
    else:
        scene, cam = ptrace.cornell(7, 320, 240)
    w, h = 320, 240
    ref = ptrace.Renderer(scene, cam, w, h, 23, False, 0, True, 8)
    ref.set_tuning(drain_threshold=-1)
    ref.add_samples(3)
    ref.add_samples(6)
    r = ptrace.Renderer(scene, cam, w, h, 23, False, 0, True, 8)
    r.set_tuning(drain_threshold=1 << 30, path_blocks_per_cu=blocks, path_shade_batch=batch, path_first_samples=first,
                 path_sort_group=group, path_trace_steps=steps)
    r.add_samples(3)
    r.add_samples(6)
    st = r.last_launch_stats()
    assert st["trace_launches"] == 0 and st["drained_paths"] == w * h
    assert st["split_launches"] == (1 if split else 0)
    (a, ga), (b, gb) = ref.read_state(), r.read_state()
    assert np.array_equal(a, b) and np.array_equal(ga, gb)
    assert ref.rays_traced() == r.rays_traced()
    sr = ref.last_launch_stats()
    assert (st["extend_rays"], st["shadow_rays"]) == (sr["extend_rays"], sr["shadow_rays"])


def test_multibounce_converges_to_reference_16384spp():
    """Full multi-bounce transport (MIS, the reference's path limit of 50) of
    the small benchmark room converges to the reference host path's own
    16384-spp render (tests/golden/converged_cornell.npz): a render with an
    unrelated seed reaches relMSE <= 1e-3 at 4096 spp, with relMSE falling
    about as 1/spp from 64 spp (a biased estimator would level off)."""
    import os
    from conftest import GOLDEN
    path = os.path.join(GOLDEN, "converged_cornell.npz")
    g = np.load(path)
    w, h = int(g["width"]), int(g["height"])
    ref = g["image"].astype(np.float64).reshape(-1, 3)
    scene, cam = ptrace.cornell(7, w, h)

    def relmse(x):
        return float(np.mean((np.asarray(x, np.float64) - ref) ** 2 / (ref ** 2 + 0.01)))

    errors = {}
    with ptrace.Renderer(scene, cam, w, h, seed=7, direct_light=True) as r:
        done = 0
        for target in (64, 512, 4096):
            while done < target:
                r.add_samples(min(256, target - done))
                done = r.sample_count()
            errors[target] = relmse(r.image())
    print("relMSE vs 16384-spp reference:", errors, flush=True)
    assert errors[4096] <= 1e-3
    assert errors[64] > 4 * errors[512] and errors[512] > 2 * errors[4096]
