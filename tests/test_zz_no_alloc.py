"""pt_renderer_add_samples_noalloc against pt_renderer_add_samples, through the
C ABI (needs an MI355X).

The allocation-free entry keeps its scratch (control block, traversal stack
spill) in one fixed device array per GPU that every call shares and that keeps
the last call's contents; each check runs two calls in a row, the second over
a smaller frame, and compares sums, PCG states, ray counts and images with
the allocating entry bit for bit.
"""
import numpy as np
import pytest

import ptrace
import tie_scenes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def device():
    ok, name = ptrace.device_available()
    assert ok, f"no HIP device: {name}"
    return name


def _pair(scene, cam, w, h, seed, mode, batches, max_depth=8):
    a = ptrace.Renderer(scene, cam, w, h, seed, False, mode, True, max_depth)
    b = ptrace.Renderer(scene, cam, w, h, seed, False, mode, True, max_depth)
    for n in batches:
        a.add_samples(n)
        b.add_samples_noalloc(n)
    return a, b


def _same(a, b):
    sa, ra = a.read_state()
    sb, rb = b.read_state()
    assert np.array_equal(sa, sb) and np.array_equal(ra, rb)
    assert a.rays_traced() == b.rays_traced()
    assert a.sample_count() == b.sample_count()
    assert np.array_equal(a.image(), b.image())


SCENES = {
    "icosphere": lambda w, h: ptrace.icosphere(64, 7, w, h),
    "cornell": lambda w, h: ptrace.cornell(7, w, h),
    "fine_grid": lambda w, h: tie_scenes.fine_grid_scene(64, 1),
}


@pytest.mark.parametrize("name", sorted(SCENES))
def test_noalloc_matches_add_samples_twice_second_smaller(name):
    for w, h, batches in ((320, 240, [3, 5]), (64, 48, [2])):
        print(f"{name} {w}x{h} {batches}", flush=True)
        scene, cam = SCENES[name](w, h)
        a, b = _pair(scene, cam, w, h, 11, ptrace.DEBUG_SHADED, batches)
        _same(a, b)


def test_noalloc_on_deep_hierarchy_and_benchmark_scene():
    # An unbalanced 63-level hierarchy (the 4-wide stack spills beyond LDS,
    # so fewer blocks stay resident), then the benchmark scene at a smaller size.
    for scene, cam, w, h in ((tie_scenes.chain_scene(63, 63), None, 96, 64),
                             (*ptrace.icosphere(224, 7, 192, 108), 192, 108)):
        if cam is None:
            cam = ptrace.make_camera((31.5, 0.5, 40), (31.5, 0.5, 0), (0, 1, 0), 60, 1.5, 0.0, 40)
        a, b = _pair(scene, cam, w, h, 5, ptrace.DEBUG_SHADED, [4])
        _same(a, b)


@pytest.mark.parametrize("mode", [ptrace.DEBUG_NORMALS, ptrace.DEBUG_BVH])
def test_noalloc_diagnostic_views(mode):
    for w, h in ((160, 100), (40, 25)):
        scene, cam = ptrace.cornell(3, w, h)
        a, b = _pair(scene, cam, w, h, 3, mode, [2, 3])
        _same(a, b)


def test_noalloc_mixed_with_allocating_calls_and_reset():
    scene, cam = ptrace.icosphere(40, 3, 128, 96)
    a = ptrace.Renderer(scene, cam, 128, 96, 9, False, 0, True, 8)
    b = ptrace.Renderer(scene, cam, 128, 96, 9, False, 0, True, 8)
    for n in (2, 3, 1):
        a.add_samples(n)
    b.add_samples(2)
    b.add_samples_noalloc(3)
    b.add_samples(1)
    _same(a, b)
    b.reset(cam)
    a.reset(cam)
    a.add_samples(4)
    b.add_samples_noalloc(4)
    _same(a, b)


def test_noalloc_rejects_what_it_cannot_serve():
    scene, cam = ptrace.cornell(7, 32, 20)
    r = ptrace.Renderer(scene, cam, 32, 20, 7, False, 0, True, 8)
    with pytest.raises(ValueError):
        r.add_samples_noalloc(0)
    r.set_tuning(exact_traversal=True)
    with pytest.raises(ValueError):
        r.add_samples_noalloc(2)
    brute = ptrace.Renderer(scene, cam, 32, 20, 7, True, 0, True, 8)
    with pytest.raises(ValueError):
        brute.add_samples_noalloc(2)
