"""Workloads of the host libraries for the sanitizer tests
(tests/test_sanitizers.py): each returns digests of its outputs, computed once
with the normal builds in the test process and once with the AddressSanitizer
and UndefinedBehaviorSanitizer builds (tests/native/build_sanitized.py) in a
child process that preloads the sanitizer runtime.

``python tests/sanitized_workload.py <scene|oracle|model>`` runs one workload
on the sanitized builds and prints its digests as JSON.
"""
import glob
import hashlib
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
for p in (os.path.join(ROOT, "cuda-path-tracer_amd"), os.path.join(ROOT, "oracle"), os.path.join(HERE, "native"), HERE):
    if p not in sys.path:
        sys.path.insert(0, p)

import numpy as np  # noqa: E402

GOLDEN = os.path.join(HERE, "golden")
IMAGES = sorted(glob.glob(os.path.join(GOLDEN, "images", "*")) +
                glob.glob(os.path.join(GOLDEN, "assets", "**", "*.png"), recursive=True) +
                glob.glob(os.path.join(GOLDEN, "assets", "**", "*.jpg"), recursive=True) +
                glob.glob(os.path.join(GOLDEN, "fuzz", "*")))
MODELS = sorted(glob.glob(os.path.join(GOLDEN, "assets", "models", "*", "*.obj")))


def _digest(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()[:16]


def _scene_digest(scene):
    return _digest(scene.triangles, scene.spheres, scene.nodes, scene.materials, scene.textures, scene.pixels,
                   np.array([scene.root, scene.max_depth]))


def scene_workload():
    """Scene library: OBJ/MTL import of every fixture model with both
    hierarchy builders, every fixture image and fuzz regression input, the
    SAH and reference-order builders on a soup, the synthetic scenes, PPM."""
    import ptrace
    out = {}
    for path in MODELS:
        for split in ("sah", "reference"):
            name = f"obj:{os.path.relpath(path, GOLDEN)}:{split}"
            try:
                scene, cam, info = ptrace.load_obj(path, 1.5, seed=7, split=split, rotate=(10, 30, 0),
                                                   scale=(1, -1.5, 1))
                out[name] = _scene_digest(scene) + _digest(np.asarray(cam))
            except Exception as e:  # the error message is part of the result
                out[name] = "error: " + str(e).replace(GOLDEN, "")
    for path in IMAGES:
        name = f"image:{os.path.relpath(path, GOLDEN)}"
        try:
            out[name] = _digest(ptrace.load_image(path))
        except Exception as e:
            out[name] = "error: " + str(e).replace(GOLDEN, "")
    soup, _ = ptrace.soup(3000, 5, 32, 20)
    out["soup"] = _scene_digest(soup)
    for split in ("sah", "median", "reference"):
        out[f"build:{split}"] = _scene_digest(ptrace.build_scene(soup.triangles, soup.spheres, soup.materials,
                                                                 soup.textures, soup.pixels, split=split, seed=3))
    out["cornell"] = _scene_digest(ptrace.cornell(7, 64, 40)[0])
    out["icosphere"] = _scene_digest(ptrace.icosphere(12, 7, 64, 40)[0])
    return out


def oracle_workload():
    """CPU oracle: shaded frames (MIS and not) and the five diagnostic views
    of the Cornell room, a ray query over a soup with its test counts."""
    import oracle
    import ptrace
    import tie_scenes
    out = {}
    scene, cam = ptrace.cornell(7, 24, 16)
    for mode in range(6):
        for mis in (False, True):
            r = oracle.Renderer(scene, cam, 24, 16, 7, False, mode, mis, 8)
            r.add_samples(2, threads=2)
            out[f"cornell:{mode}:{int(mis)}"] = _digest(r.sums(), r.rng_state(), np.array([r.rays_traced()]))
    soup, _ = ptrace.soup(2000, 3, 24, 16)
    res, boxes, prims = oracle.query(soup, tie_scenes.random_rays(soup, 3000, 2))
    out["query"] = _digest(res, boxes, prims)
    return out


def model_workload(sanitized):
    """CPU model of the traversal: the constructed order-sensitive scenes and
    an unbalanced hierarchy in all three modes."""
    import tie_scenes
    import wide_model as wm
    out = {}
    cases = {"grid": (tie_scenes.grid_scene(6, 1e-6, 1), tie_scenes.edge_rays(6, 3000, 1)),
             "spheres": (tie_scenes.sphere_scene(20, 2), None),
             "chain": (tie_scenes.chain_scene(17, 17), None)}
    for name, (scene, rays) in cases.items():
        if rays is None:
            rays = tie_scenes.random_rays(scene, 3000, 4)
        for mode in (wm.WIDE, wm.WIDE_NO_FALLBACK, wm.REFERENCE_ORDER):
            hits, stats, overflows = wm.trace(scene, rays, mode, sanitized=sanitized)
            out[f"{name}:{mode}"] = _digest(hits, stats, np.array([overflows]))
    return out


def run(which, sanitized):
    """Digests of one workload; `sanitized` points the libraries at the
    sanitizer builds (only in a process that preloads the runtime)."""
    if sanitized:
        import oracle
        import ptrace
        native = os.path.join(HERE, "native")
        ptrace.SCENE_LIBRARY_PATH = os.path.join(native, "libptscene_asan.so")
        oracle.ORACLE_PATH = os.path.join(native, "liboracle_asan.so")
    if which == "scene":
        return scene_workload()
    if which == "oracle":
        return oracle_workload()
    if which == "model":
        return model_workload(sanitized)
    raise ValueError(which)


if __name__ == "__main__":
    print(json.dumps(run(sys.argv[1], sanitized=True)))
