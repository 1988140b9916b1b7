"""CPU-only checks of the native libraries: exported symbols, host-side
validation through the C ABI, and the BVH builder's flat output.  No GPU
compute is launched here."""
import ctypes
import os
import re

import numpy as np
import pytest

import oracle
import ptrace
from conftest import ROOT, golden


def _declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pts?_[a-z_0-9]+)\s*\(", text)) - {"pt_status"})


@pytest.mark.parametrize("header,path", [("pt_render.h", ptrace.LIBRARY_PATH), ("pt_scene.h", ptrace.SCENE_LIBRARY_PATH)])
def test_library_exports_every_declared_symbol(header, path):
    assert os.path.exists(path), f"{path} not built"
    lib = ctypes.CDLL(path)
    names = _declared(header)
    assert names, header
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_ctypes_table_covers_the_header():
    assert set(_declared("pt_render.h")) == set(ptrace.PT_FUNCTIONS)


def test_record_layouts_match_the_reference_sizes():
    for dt, size in ((ptrace.TRIANGLE, 116), (ptrace.SPHERE, 20), (ptrace.NODE, 40), (ptrace.MATERIAL, 28),
                     (ptrace.TEXTURE, 60), (ptrace.CAMERA, 84), (ptrace.HIT, 48), (ptrace.TRACE_RESULT, 56),
                     (ptrace.RAY, 28)):
        assert dt.itemsize == size


def test_device_probe_reports_cleanly():
    ok, text = ptrace.device_available()
    assert isinstance(ok, bool) and text


def test_boundary_validation_runs_before_the_device():
    # cuda_renderer.cu:177-179: scene, camera and mode are validated first and
    # map to runtime_error / invalid_argument.
    scene, cam, _ = golden("asset_cube.npz")
    broken = ptrace.FlatScene(scene.triangles.copy(), scene.spheres, scene.nodes.copy(), scene.materials,
                              scene.textures, scene.pixels, scene.root, scene.max_depth)
    broken.nodes["left"][0] = 999
    with pytest.raises(RuntimeError, match="invalid flat scene data"):
        ptrace.Renderer(broken, cam, 8, 8, 1)
    bad_cam = np.array(cam, copy=True)
    bad_cam["time1"] = -5
    with pytest.raises(ptrace.InvalidArgument, match="lens/shutter"):
        ptrace.Renderer(scene, bad_cam, 8, 8, 1)
    with pytest.raises(ptrace.InvalidArgument, match="debug mode"):
        ptrace.Renderer(scene, cam, 8, 8, 1, debug=9)
    with pytest.raises(ptrace.InvalidArgument, match="pixel count"):
        ptrace.Renderer(scene, cam, 0, 8, 1)
    with pytest.raises(ptrace.InvalidArgument, match="pixel count"):
        ptrace.Renderer(scene, cam, 65536, 65536, 1)
    with pytest.raises(RuntimeError, match="invalid flat scene data"):
        ptrace.query_rays(broken, np.zeros(1, ptrace.RAY))


def test_host_buffer_arguments_are_checked_before_the_device():
    import ctypes
    L = ptrace.lib()
    out = ctypes.c_void_p()
    assert L.pt_host_alloc(0, ctypes.byref(out)) == ptrace.PT_EINVAL and out.value is None
    assert L.pt_host_alloc(16, None) == ptrace.PT_EINVAL
    assert L.pt_host_free(None) == ptrace.PT_OK


def test_device_bvh_arguments_are_checked_before_the_device():
    import ctypes
    L = ptrace.lib()
    tri = np.zeros(2, ptrace.TRIANGLE)
    tri["v1"], tri["v2"] = (1, 0, 0), (0, 1, 0)
    nodes = np.zeros(2, ptrace.NODE)  # 2n - 1 = 3 needed
    root, depth, ms = ctypes.c_int32(7), ctypes.c_int32(7), ctypes.c_float(7)
    out = (ctypes.byref(root), ctypes.byref(depth), ctypes.byref(ms))
    p = lambda a: ctypes.c_void_p(a.ctypes.data)
    assert L.pt_build_bvh(p(tri), 2, None, 0, p(nodes), 2, *out) == ptrace.PT_EINVAL
    assert "2n - 1" in L.pt_last_error().decode() and root.value == -1 and depth.value == 0 and ms.value == 0
    assert L.pt_build_bvh(None, 2, None, 0, p(nodes), 2, *out) == ptrace.PT_EINVAL
    assert L.pt_build_bvh(p(tri), -1, None, 0, p(nodes), 2, *out) == ptrace.PT_EINVAL
    assert L.pt_build_bvh(p(tri), 2, None, 0, p(nodes), 3, None, *out[1:]) == ptrace.PT_EINVAL
    bad = tri.copy()
    bad["v2"][1, 0] = np.inf
    assert L.pt_build_bvh(p(bad), 2, None, 0, p(nodes), 3, *out) == ptrace.PT_ESCENE
    sph = np.zeros(1, ptrace.SPHERE)  # radius 0
    assert L.pt_build_bvh(None, 0, p(sph), 1, p(nodes), 3, *out) == ptrace.PT_ESCENE
    # No primitives: the empty hierarchy, without touching the device.
    assert L.pt_build_bvh(None, 0, None, 0, None, 0, *out) == ptrace.PT_OK
    assert root.value == -1 and depth.value == 0


@pytest.mark.parametrize("fixture", ["asset_cube.npz", "asset_sphere.npz", "asset_triangle.npz", "cornell.npz",
                                     "pruning.npz", "spheres.npz"])
def test_c_abi_validation_agrees_with_oracle(fixture):
    scene, _, _ = golden(fixture)
    assert oracle.validate(scene)
    rng = np.random.default_rng(5)
    for trial in range(12):
        s = ptrace.FlatScene(scene.triangles.copy(), scene.spheres.copy(), scene.nodes.copy(), scene.materials.copy(),
                             scene.textures.copy(), scene.pixels.copy(), scene.root, scene.max_depth)
        k = int(rng.integers(len(s.nodes)))
        field = ["left", "right", "primitive", "type"][trial % 4]
        s.nodes[field][k] = int(rng.integers(-2, len(s.nodes) + 2))
        expect = oracle.validate(s)
        try:
            ptrace.Renderer(s, golden(fixture)[1] if golden(fixture)[1] is not None else _cam(), 4, 4, 1)
            host_ok = True
        except RuntimeError as e:
            host_ok = "invalid flat scene data" not in str(e)
        if not expect:
            assert not host_ok, (fixture, field, k)


def _cam():
    return ptrace.make_camera((0, 0, 5), (0, 0, 0), (0, 1, 0), 40, 1, 0, 5)


@pytest.mark.parametrize("maker", ["cornell", "soup", "icosphere"])
def test_builder_emits_valid_depth_first_flat_scenes(maker):
    if maker == "cornell":
        scene, cam = ptrace.cornell(3, 32, 20)
    elif maker == "soup":
        scene, cam = ptrace.soup(5000, 9, 32, 20)
    else:
        scene, cam = ptrace.icosphere(12, 5, 32, 20)
    assert oracle.validate(scene) and scene.max_depth <= 64
    n = scene.nodes
    branch = n["type"] == ptrace.BRANCH
    # depth-first preorder: a branch's left child follows it
    assert (n["left"][branch] == np.nonzero(branch)[0] + 1).all()
    # primitives renumbered in leaf order
    tri_leaf = n["primitive"][n["type"] == ptrace.TRIANGLE_PRIM]
    assert np.array_equal(tri_leaf, np.arange(len(scene.triangles)))
    # leaf bounds are the reference's padded triangle boxes (triangle.h:96-109)
    leaves = np.nonzero(n["type"] == ptrace.TRIANGLE_PRIM)[0][:200]
    for i in leaves:
        t = scene.triangles[n["primitive"][i]]
        v = np.stack([t["v0"], t["v1"], t["v2"]])
        lo = np.nextafter(v.min(axis=0) - np.float32(1e-4), np.float32(-np.inf))
        hi = np.nextafter(v.max(axis=0) + np.float32(1e-4), np.float32(np.inf))
        assert np.array_equal(n["minimum"][i], lo) and np.array_equal(n["maximum"][i], hi)
    lights = oracle.scene_lights(scene)
    assert len(lights) >= 1


def test_icosphere_config_size():
    # 20*n^2 triangles of the displaced sphere plus 2 floor and 8 light triangles
    scene, _ = ptrace.icosphere(16, 7, 32, 20)
    assert len(scene.triangles) == 20 * 16 * 16 + 10
    assert (scene.triangles["smooth"][:20 * 16 * 16].sum() + scene.triangles["smooth"].sum()) > 0


def test_camera_basis_matches_reference_constructor():
    if not oracle.ref_available():
        pytest.skip("compiled reference not present")
    L = oracle.ref()
    args = ((0.3, 1.2, 4.0), (0.0, 0.9, -0.2), (0, 1, 0), 37.0, 1.6, 0.1, 3.7, 0.2, 0.9)
    want = np.zeros((), ptrace.CAMERA)
    f = [np.asarray(a, np.float32) for a in args[:3]]
    L.ref_camera(f[0].ctypes.data, f[1].ctypes.data, f[2].ctypes.data, *[ctypes.c_float(a) for a in args[3:]],
                 want.ctypes.data)
    got = ptrace.make_camera(*args)
    assert got.tobytes() == want.tobytes()
