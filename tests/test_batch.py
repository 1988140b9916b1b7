"""Batched rendering (pt_batch_*): many views of many scenes in one
graph-captured bounce loop.  Every view must equal what a single renderer
makes of the same scene, camera, size and seed, bit for bit."""
import os

import numpy as np
import pytest

import oracle
import ptrace
from conftest import golden

THREADS = min(16, len(os.sched_getaffinity(0)))


def _scenes():
    a, ca = ptrace.cornell(3, 40, 30)
    b, cb = ptrace.soup(6000, 2, 48, 32)
    c, cc = ptrace.icosphere(20, 5, 36, 36)
    d, cd, _ = golden("asset_sphere.npz")
    return [a, b, c, d], [ca, cb, cc, cd]


@pytest.mark.gpu
@pytest.mark.parametrize("depth,direct", [(6, True), (8, False)])
def test_batch_views_match_single_renderers(depth, direct):
    scenes, cams = _scenes()
    views = [(0, cams[0], 40, 30, 7), (1, cams[1], 48, 32, 9), (2, cams[2], 36, 36, 3), (3, cams[3], 20, 14, 5),
             (0, cams[0], 24, 18, 11)]
    batch = ptrace.Batch(scenes, views, direct_light=direct, max_depth=depth)
    batch.add_samples(3)
    batch.add_samples(2)
    total = 0
    for k, (si, cam, w, h, seed) in enumerate(views):
        r = ptrace.Renderer(scenes[si], cam, w, h, seed, False, ptrace.DEBUG_SHADED, direct, depth)
        r.add_samples(3)
        r.add_samples(2)
        assert np.array_equal(batch.image(k), r.image()), k
        total += r.rays_traced()
    st = batch.stats()
    assert st["samples"] == 5 and st["rays"] == total


@pytest.mark.gpu
def test_batch_of_benchmark_sized_scenes_matches_single_renderers_and_oracle():
    # The batched configuration's range: clustered triangle soups of 10k to
    # 2M triangles (the batch benchmark's generator), uneven resolutions from
    # 256x256 to 1280x720, one scene seen twice, path depth 6 with MIS.
    # Every view equals a single renderer bit for bit (and the batch's ray
    # count their sum); every single renderer equals the oracle on the whole
    # frame: per-pixel sums, PCG states and the exact ray count.
    sizes = [10_000, 60_000, 300_000, 1_000_000, 2_000_000]
    shapes = [(256, 256), (640, 480), (1024, 576), (800, 600), (1280, 720)]
    built = [ptrace.soup(n, 100 + k, w, h) for k, (n, (w, h)) in enumerate(zip(sizes, shapes))]
    scenes = [b[0] for b in built]
    views = [(k, built[k][1], w, h, 31 + 7 * k) for k, (w, h) in enumerate(shapes)]
    # the 2M-triangle scene again, from its generator's camera for 512x384
    views.append((4, ptrace.soup(16, 104, 512, 384)[1], 512, 384, 73))
    batch = ptrace.Batch(scenes, views, direct_light=True, max_depth=6)
    batch.add_samples(2)
    batch.add_samples(1)
    total = 0
    for k, (si, cam, w, h, seed) in enumerate(views):
        print(f"view {k}: {len(scenes[si].triangles)} triangles, {w}x{h}", flush=True)
        r = ptrace.Renderer(scenes[si], cam, w, h, seed, False, ptrace.DEBUG_SHADED, True, 6)
        r.add_samples(2)
        r.add_samples(1)
        assert np.array_equal(batch.image(k), r.image()), k
        total += r.rays_traced()
        o = oracle.Renderer(scenes[si], cam, w, h, seed, False, 0, True, 6)
        o.add_samples(2, threads=THREADS)
        o.add_samples(1, threads=THREADS)
        sums, rng = r.read_state()
        assert np.array_equal(sums, o.sums()), k
        assert np.array_equal(rng, o.rng_state()), k
        assert r.rays_traced() == o.rays_traced(), k
    st = batch.stats()
    assert st["samples"] == 3 and st["rays"] == total


@pytest.mark.gpu
def test_batch_split_path_lanes_match_single_renderers():
    # 68544 pixels on 65536 lanes (one block per compute unit): the batch's
    # path-lane launch of 6 samples runs in two passes (4 samples of every
    # pixel, then the rest with groups of 50 pixels reordered by cost, the
    # last group partial); the single renderers run one pass.  Images and ray
    # counts must be equal.
    scenes, cams = _scenes()
    views = [(0, cams[0], 200, 150, 7), (2, cams[2], 180, 180, 3), (1, cams[1], 96, 64, 9)]
    batch = ptrace.Batch(scenes[:3], views, direct_light=True, max_depth=8)
    batch.set_tuning(path_blocks_per_cu=1, path_sort_group=50, path_first_samples=4)
    batch.add_samples(3)
    batch.add_samples(6)
    assert batch.last_launch_stats()["split_launches"] == 1
    total = 0
    for k, (si, cam, w, h, seed) in enumerate(views):
        r = ptrace.Renderer(scenes[si], cam, w, h, seed, False, ptrace.DEBUG_SHADED, True, 8)
        r.set_tuning(path_first_samples=-1)
        r.add_samples(3)
        r.add_samples(6)
        assert np.array_equal(batch.image(k), r.image()), k
        total += r.rays_traced()
    assert batch.stats()["rays"] == total


@pytest.mark.gpu
def test_batch_tuning_and_autotuner_keep_images():
    scenes, cams = _scenes()
    views = [(0, cams[0], 40, 30, 7), (1, cams[1], 48, 32, 9), (2, cams[2], 36, 36, 3)]
    batch = ptrace.Batch(scenes[:3], views, direct_light=True, max_depth=6)
    batch.add_samples(4)
    ref = [batch.image(k) for k in range(3)]
    best, rate, trials = batch.autotune(probe_samples=1, passes=1)
    assert rate > 0 and len(trials) >= 10 and best["trace_block"] in (64, 128, 192, 256)
    assert batch.stats()["samples"] == 0
    batch.add_samples(4)
    for k in range(3):
        assert np.array_equal(batch.image(k), ref[k]), k
    batch.reset()
    batch.set_tuning(trace_block=64, shade_block=128, check_interval=2, drain_threshold=-1, lane_refill=8)
    batch.add_samples(4)
    for k in range(3):
        assert np.array_equal(batch.image(k), ref[k]), k
    rays = batch.stats()["rays"]
    # Queries handed on between launches after one step on empty queues.
    batch.reset()
    batch.set_tuning(handoff_steps=1)
    batch.add_samples(4)
    for k in range(3):
        assert np.array_equal(batch.image(k), ref[k]), k
    assert batch.stats()["rays"] == rays


@pytest.mark.gpu
def test_batch_graph_follows_reallocated_buffers():
    # Raising the drain threshold between calls grows the traversal spill
    # region past what the captured bounce-loop graph was made with; the
    # graph must be captured again (it holds buffer addresses by value).
    a, ca = ptrace.soup(20000, 4, 128, 128)
    b, cb = ptrace.icosphere(24, 2, 128, 128)
    views = [(0, ca, 128, 128, 3), (1, cb, 128, 128, 4)]
    ref = ptrace.Batch([a, b], views, direct_light=True, max_depth=6)
    ref.set_tuning(drain_threshold=1000)
    ref.add_samples(2)
    ref.add_samples(2)
    batch = ptrace.Batch([a, b], views, direct_light=True, max_depth=6)
    batch.set_tuning(blocks_per_cu=1, trace_block=64, drain_threshold=1000)
    batch.add_samples(2)
    batch.set_tuning(drain_threshold=20000)
    batch.add_samples(2)
    for k in range(2):
        assert np.array_equal(batch.image(k), ref.image(k)), k


def test_batch_argument_checks():
    scenes, cams = _scenes()
    with pytest.raises(ptrace.InvalidArgument):
        ptrace.Batch(scenes[:1], [(1, cams[0], 8, 8, 1)])
    with pytest.raises(ptrace.InvalidArgument):
        ptrace.Batch(scenes[:1], [(0, cams[0], 0, 8, 1)])
