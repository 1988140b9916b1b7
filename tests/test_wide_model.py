"""CPU model of the shaded path's 4-wide traversal (tests/native/wide_model.hip,
built from the renderer's own traversal and layout sources) against the
oracle's reference-order ray query (oracle/pt_oracle.c, which follows
src/renderer/transport.h:43-85).

- The reference-order traversal over the renderer's binary layout equals the
  oracle on every ray.
- The 4-wide nearest-first traversal with its reference-order fallback equals
  the oracle on every ray, on ordinary scenes and on constructed ones where
  the order decides the closest hit (coplanar axis-aligned triangles with
  tight boxes hit through their shared edges; tight sphere boxes and grazing
  rays).  Without the fallback it differs on the constructed scenes, which
  shows that the cases exist and that the fallback is what handles them.
- The traversal stack never holds more than 3 * (wide depth) + 1 entries,
  the capacity the renderer sizes it for, and never overflows.
"""
import os
import sys

import numpy as np
import pytest

import oracle
import ptrace
import tie_scenes

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "native"))
import wide_model as wm  # noqa: E402


def _oracle(scene, rays):
    res, _, _ = oracle.query(scene, rays)
    return tie_scenes.oracle_hits(res)


def _same(model_hits, ref_hits):
    """Per ray: same closest distance (bits) and same primitive."""
    return (model_hits[:, 0] == ref_hits[:, 0]) & (model_hits[:, 3] == ref_hits[:, 3])


def _check_stack(scene, stats, overflows):
    info = wm.layout_info(scene)
    assert overflows == 0
    assert stats[:, 2].max() <= info["wide_capacity"]
    # without a retrace the 4-wide stack stays within 3 * depth + 1
    plain = stats[:, 3] == 0
    if plain.any():
        assert stats[plain, 2].max() <= 3 * info["wide_depth"] + 1


ORDINARY = {
    "cornell": lambda: ptrace.cornell(7, 32, 20)[0],
    "icosphere": lambda: ptrace.icosphere(24, 7, 32, 20)[0],
    "soup": lambda: ptrace.soup(20000, 3, 32, 20)[0],
}


@pytest.mark.parametrize("name", sorted(ORDINARY))
def test_model_matches_oracle_on_ordinary_scenes(name):
    scene = ORDINARY[name]()
    rays = tie_scenes.random_rays(scene, 50000, 11)
    ref = _oracle(scene, rays)
    for mode in (wm.WIDE, wm.REFERENCE_ORDER):
        hits, stats, overflows = wm.trace(scene, rays, mode)
        assert _same(hits, ref).all(), (name, mode)
        if mode == wm.WIDE:
            _check_stack(scene, stats, overflows)
        else:
            assert overflows == 0 and stats[:, 2].max() <= scene.max_depth


@pytest.mark.parametrize("seed", range(6))
def test_fallback_makes_coplanar_edge_ties_follow_the_reference(seed):
    for pad in (1e-7, 1e-6):
        scene = tie_scenes.grid_scene(6, pad, seed)
        oracle.validate(scene)
        rays = tie_scenes.edge_rays(6, 20000, seed)
        ref = _oracle(scene, rays)
        hits, stats, overflows = wm.trace(scene, rays, wm.WIDE)
        assert _same(hits, ref).all()
        # the grid's barycentric texture coordinates are the hit's u, v exactly
        assert (hits[:, 1:3] == ref[:, 1:3]).all()
        _check_stack(scene, stats, overflows)
        assert stats[:, 3].any()  # the fallback ran
        order, _, _ = wm.trace(scene, rays, wm.REFERENCE_ORDER)
        assert _same(order, ref).all()


def test_without_fallback_coplanar_edge_ties_differ():
    # The cases the fallback handles exist: over these seeds the traversal
    # without it picks another of two tied triangles than the reference on a
    # few rays in 10,000.
    differ = 0
    for seed in range(6):
        for pad in (1e-7, 1e-6):
            scene = tie_scenes.grid_scene(6, pad, seed)
            rays = tie_scenes.edge_rays(6, 20000, seed)
            bare, _, _ = wm.trace(scene, rays, wm.WIDE_NO_FALLBACK)
            differ += int((~_same(bare, _oracle(scene, rays))).sum())
    assert differ > 0


def test_camera_rays_on_shared_edges():
    # The render-side case: camera rays of fine_grid_scene land exactly on the
    # grid's edges; without the fallback some pick another tied triangle.
    differ = 0
    for seed in range(3):
        scene, cam = tie_scenes.fine_grid_scene(64, seed)
        oracle.validate(scene)
        rays = tie_scenes.camera_rays(cam, 96, 96, 50000, seed)
        ref = _oracle(scene, rays)
        assert (ref[:, 3] >= 0).all()
        hits, stats, overflows = wm.trace(scene, rays, wm.WIDE)
        assert _same(hits, ref).all()
        assert stats[:, 3].sum() > 100
        _check_stack(scene, stats, overflows)
        bare, _, _ = wm.trace(scene, rays, wm.WIDE_NO_FALLBACK)
        differ += int((~_same(bare, ref)).sum())
    assert differ > 0


def test_flat_leaf_boxes_do_not_hide_hits():
    # pad 0: branch boxes of coplanar triangles are flat too, so the reference
    # finds nothing through them; the 4-wide traversal must agree (a flat box
    # is never the only test between a ray and a hit it reports).
    scene = tie_scenes.grid_scene(6, 0.0, 3)
    rays = tie_scenes.edge_rays(6, 20000, 3)
    ref = _oracle(scene, rays)
    hits, _, _ = wm.trace(scene, rays, wm.WIDE)
    assert _same(hits, ref).all()


@pytest.mark.parametrize("seed", range(4))
def test_tight_sphere_boxes_and_grazing_rays(seed):
    scene = tie_scenes.sphere_scene(40, seed)
    oracle.validate(scene)
    rays = np.concatenate([tie_scenes.grazing_rays(scene, 20000, seed), tie_scenes.random_rays(scene, 5000, seed)])
    ref = _oracle(scene, rays)
    assert (ref[:, 3] >= 0).sum() > 5000
    hits, stats, overflows = wm.trace(scene, rays, wm.WIDE)
    assert _same(hits, ref).all()
    _check_stack(scene, stats, overflows)


def test_any_hit_matches_reference_order():
    scene = tie_scenes.grid_scene(6, 1e-6, 1)
    rays = tie_scenes.edge_rays(6, 20000, 1)
    ref = _oracle(scene, rays)
    t = ref[:, 0].view(np.float32)
    rng = np.random.default_rng(1)
    # upper bounds just below, at and above the closest hit, and unbounded
    high = np.where(ref[:, 3] >= 0, t * rng.choice([0.999, 1.0, 1.001], len(t)).astype(np.float32),
                    np.float32(3e38)).astype(np.float32)
    a, _, _ = wm.trace(scene, rays, wm.WIDE, any_hit=True, high=high)
    b, _, _ = wm.trace(scene, rays, wm.REFERENCE_ORDER, any_hit=True, high=high)
    assert ((a[:, 3] >= 0) == (b[:, 3] >= 0)).all()


@pytest.mark.parametrize("length", [3, 17, 40, 63])
def test_stack_bound_on_unbalanced_hierarchies(length):
    scene = tie_scenes.chain_scene(length, length)
    oracle.validate(scene)
    info = wm.layout_info(scene)
    rays = tie_scenes.random_rays(scene, 20000, length)
    ref = _oracle(scene, rays)
    hits, stats, overflows = wm.trace(scene, rays, wm.WIDE)
    assert _same(hits, ref).all()
    _check_stack(scene, stats, overflows)
    assert info["wide_capacity"] >= max(3 * info["wide_depth"] + 1, info["binary_capacity"])
