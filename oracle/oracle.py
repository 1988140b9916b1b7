"""ctypes access to the CPU oracle (liboracle.so) and, when it was built in
this container, the compiled reference path (oracle/_ref/libptref.so).

TEST INFRASTRUCTURE ONLY: used by tests/, the smoke entry and the benchmark's
CPU-baseline leg as the checker.  The product never imports it.

Scenes are duck-typed: any object with the flat arrays (triangles, spheres,
nodes, materials, textures, pixels as numpy structured arrays with the
reference's record layouts) plus root and max_depth.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libptref.so")

SIZES = {"triangles": 116, "spheres": 20, "nodes": 40, "materials": 28, "textures": 60, "pixels": 1}


class _Scene(ctypes.Structure):
    _fields_ = [("triangles", ctypes.c_void_p), ("triangle_count", ctypes.c_int64),
                ("spheres", ctypes.c_void_p), ("sphere_count", ctypes.c_int64),
                ("nodes", ctypes.c_void_p), ("node_count", ctypes.c_int64),
                ("materials", ctypes.c_void_p), ("material_count", ctypes.c_int64),
                ("textures", ctypes.c_void_p), ("texture_count", ctypes.c_int64),
                ("pixels", ctypes.c_void_p), ("pixel_bytes", ctypes.c_int64),
                ("root", ctypes.c_int32), ("max_depth", ctypes.c_int32)]


class _Rng(ctypes.Structure):
    _fields_ = [("state", ctypes.c_uint64), ("increment", ctypes.c_uint64)]


class _Vec3(ctypes.Structure):
    _fields_ = [("x", ctypes.c_float), ("y", ctypes.c_float), ("z", ctypes.c_float)]


class _Ray(ctypes.Structure):
    _fields_ = [("origin", _Vec3), ("direction", _Vec3), ("time", ctypes.c_float)]


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data) if a.size else ctypes.c_void_p(0)


def _arrays(scene):
    out = {}
    for name, size in SIZES.items():
        a = np.ascontiguousarray(getattr(scene, name))
        assert a.dtype.itemsize == size, (name, a.dtype.itemsize)
        out[name] = a
    return out


def scene_struct(scene):
    """(struct, keepalive) for one scene."""
    a = _arrays(scene)
    s = _Scene(_ptr(a["triangles"]), len(a["triangles"]), _ptr(a["spheres"]), len(a["spheres"]),
               _ptr(a["nodes"]), len(a["nodes"]), _ptr(a["materials"]), len(a["materials"]),
               _ptr(a["textures"]), len(a["textures"]), _ptr(a["pixels"]), len(a["pixels"]),
               int(scene.root), int(scene.max_depth))
    return s, a


_ORACLE = None
_REF = None


def build():
    """Compile liboracle.so, and oracle/_ref when the reference tree is present."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    subprocess.run(["bash", os.path.join(HERE, "build_ref.sh")], check=True, stdout=subprocess.DEVNULL)


def oracle():
    global _ORACLE
    if _ORACLE is None:
        if not os.path.exists(ORACLE_PATH):
            raise RuntimeError(f"oracle not built: {ORACLE_PATH}")
        L = ctypes.CDLL(ORACLE_PATH)
        L.po_rng_next.restype = ctypes.c_uint32
        L.po_rng_uniform.restype = ctypes.c_float
        L.po_rng_seed.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint64]
        L.po_sample_ball.restype = _Vec3
        L.po_sample_disk.restype = _Vec3
        L.po_mis_weight.restype = ctypes.c_float
        L.po_mis_weight.argtypes = [ctypes.c_float, ctypes.c_float]
        L.po_cosine_direction.restype = _Vec3
        L.po_cosine_direction.argtypes = [_Vec3, ctypes.c_void_p]
        L.po_id_color.restype = _Vec3
        L.po_id_color.argtypes = [ctypes.c_int32]
        L.po_trace_box.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_float, ctypes.c_float]
        L.po_triangle_hit.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_float, ctypes.c_float, ctypes.c_void_p]
        L.po_sphere_hit.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_float, ctypes.c_float, ctypes.c_void_p]
        L.po_light_direction_pdf.restype = ctypes.c_float
        L.po_light_direction_pdf.argtypes = [ctypes.c_float, _Vec3, _Vec3, _Vec3]
        L.po_pack_channel.argtypes = [ctypes.c_float]
        L.po_scene_lights.restype = ctypes.c_int64
        L.po_scene_lights.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64]
        L.po_validate_scene.argtypes = [ctypes.c_void_p]
        L.po_validate_camera.argtypes = [ctypes.c_void_p]
        L.po_query_rays.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_int,
                                    ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        L.po_trace_color.restype = _Vec3
        L.po_trace_color.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, _Ray, ctypes.c_void_p,
                                     ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                     ctypes.c_void_p]
        L.po_renderer_create.restype = ctypes.c_void_p
        L.po_renderer_create.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_uint64,
                                         ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.po_renderer_destroy.argtypes = [ctypes.c_void_p]
        L.po_renderer_reset.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        L.po_renderer_add_samples.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
        L.po_renderer_add_samples_rows.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        for name in ("po_renderer_image", "po_renderer_sum", "po_renderer_rng", "po_renderer_rgba"):
            getattr(L, name).argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        L.po_renderer_sample_count.argtypes = [ctypes.c_void_p]
        L.po_renderer_rays.restype = ctypes.c_uint64
        L.po_renderer_rays.argtypes = [ctypes.c_void_p]
        _ORACLE = L
    return _ORACLE


def ref_available():
    return os.path.exists(REF_PATH)


def ref():
    global _REF
    if _REF is None:
        if not ref_available():
            raise RuntimeError("compiled reference not built (oracle/_ref)")
        L = ctypes.CDLL(REF_PATH)
        L.ref_validate.argtypes = [ctypes.c_void_p]
        L.ref_scene_lights.restype = ctypes.c_int64
        L.ref_scene_lights.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64]
        L.ref_query.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_int, ctypes.c_void_p,
                                ctypes.c_void_p, ctypes.c_void_p]
        L.ref_flat_hit_world.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p]
        L.ref_rng_stream.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p]
        L.ref_samplers.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p,
                                   ctypes.c_void_p]
        L.ref_trace_color.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int,
                                      ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        L.ref_render.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_uint64,
                                 ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                                 ctypes.c_void_p]
        L.ref_render_rows.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_uint64,
                                      ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                      ctypes.c_void_p, ctypes.c_void_p]
        L.ref_display_pixels.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p]
        L.ref_build_flat.restype = ctypes.c_void_p
        L.ref_build_flat.argtypes = [ctypes.c_void_p, ctypes.c_long]
        L.ref_load_image.argtypes = [ctypes.c_char_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_long]
        L.ref_load_obj.restype = ctypes.c_void_p
        L.ref_load_obj.argtypes = [ctypes.c_char_p, ctypes.c_float, ctypes.c_long, ctypes.c_void_p, ctypes.c_void_p,
                                   ctypes.c_void_p, ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float]
        L.ref_sphere_scene.restype = ctypes.c_void_p
        L.ref_sphere_scene.argtypes = [ctypes.c_long]
        L.ref_triangle_light_scene.restype = ctypes.c_void_p
        L.ref_triangle_light_scene.argtypes = [ctypes.c_int]
        L.ref_camera.argtypes = [ctypes.c_void_p] * 3 + [ctypes.c_float] * 6 + [ctypes.c_void_p]
        L.ref_scene_counts.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        L.ref_scene_copy.argtypes = [ctypes.c_void_p] * 8
        L.ref_scene_free.argtypes = [ctypes.c_void_p]
        _REF = L
    return _REF


# ------------------------------------------------------------ oracle calls
RESULT_SIZE = 56


def query(scene, rays, brute=False):
    """(results bytes as structured array of the caller's dtype, boxes, prims)."""
    L = oracle()
    s, keep = scene_struct(scene)
    rays = np.ascontiguousarray(rays)
    assert rays.dtype.itemsize == 28
    out = np.zeros(len(rays), dtype=np.dtype((np.void, RESULT_SIZE)))
    boxes = np.zeros(len(rays), np.int32)
    prims = np.zeros(len(rays), np.int32)
    L.po_query_rays(ctypes.byref(s), _ptr(rays), len(rays), int(bool(brute)), _ptr(out), _ptr(boxes), _ptr(prims))
    return out, boxes, prims


def scene_lights(scene):
    L = oracle()
    s, keep = scene_struct(scene)
    n = L.po_scene_lights(ctypes.byref(s), None, 0)
    if n < 0:
        raise ValueError("light area exceeds range")
    out = np.zeros(n, np.dtype([("type", "<i4"), ("primitive", "<i4"), ("cumulative_area", "<f4")]))
    L.po_scene_lights(ctypes.byref(s), _ptr(out), n)
    return out


def validate(scene):
    s, keep = scene_struct(scene)
    return oracle().po_validate_scene(ctypes.byref(s)) == 0


def rng_seed(seed, pixel):
    r = _Rng()
    oracle().po_rng_seed(ctypes.byref(r), seed, pixel)
    return r


def rng_stream(seed, pixel, n):
    L = oracle()
    r = rng_seed(seed, pixel)
    return np.array([L.po_rng_next(ctypes.byref(r)) for _ in range(n)], np.uint32)


class Renderer:
    """portable_renderer (tests/support/portable_renderer.h:11-57) on the oracle."""

    def __init__(self, scene, camera, width, height, seed, brute=False, debug=0, direct_light=False, max_depth=50):
        self.L = oracle()
        self._s, self._keep = scene_struct(scene)
        self._cam = np.ascontiguousarray(np.asarray(camera).reshape(()))
        self.width, self.height = int(width), int(height)
        self._h = self.L.po_renderer_create(ctypes.byref(self._s), _ptr(self._cam), self.width, self.height,
                                            int(seed) & (2**64 - 1), int(bool(brute)), int(debug),
                                            int(bool(direct_light)), int(max_depth))
        if not self._h:
            raise ValueError("invalid oracle renderer arguments")

    def __del__(self):
        if getattr(self, "_h", None):
            self.L.po_renderer_destroy(self._h)
            self._h = None

    @property
    def pixels(self):
        return self.width * self.height

    def reset(self, camera):
        self._cam = np.ascontiguousarray(np.asarray(camera).reshape(()))
        if self.L.po_renderer_reset(self._h, _ptr(self._cam)) != 0:
            raise ValueError("invalid camera")

    def add_samples(self, count, threads=1, rows=None):
        if rows is None:
            rc = self.L.po_renderer_add_samples(self._h, int(count), int(threads))
        else:
            rc = self.L.po_renderer_add_samples_rows(self._h, int(count), int(threads), int(rows[0]), int(rows[1]))
        if rc == -1:
            raise ValueError("invalid accumulated sample count")
        if rc == -2:
            raise RuntimeError("portable traversal stack overflow")

    def image(self):
        out = np.empty((self.pixels, 3), np.float32)
        self.L.po_renderer_image(self._h, _ptr(out))
        return out

    def sums(self):
        out = np.empty((self.pixels, 3), np.float32)
        self.L.po_renderer_sum(self._h, _ptr(out))
        return out

    def rng_state(self):
        out = np.empty(self.pixels, np.uint64)
        self.L.po_renderer_rng(self._h, _ptr(out))
        return out

    def rgba_image(self):
        out = np.empty(self.pixels * 4, np.uint8)
        self.L.po_renderer_rgba(self._h, _ptr(out))
        return out

    def sample_count(self):
        return self.L.po_renderer_sample_count(self._h)

    def rays_traced(self):
        return self.L.po_renderer_rays(self._h)


# ------------------------------------------------------------ reference calls
def ref_query(scene, rays, brute=False):
    L = ref()
    s, keep = scene_struct(scene)
    rays = np.ascontiguousarray(rays)
    out = np.zeros(len(rays), dtype=np.dtype((np.void, RESULT_SIZE)))
    boxes = np.zeros(len(rays), np.int32)
    prims = np.zeros(len(rays), np.int32)
    L.ref_query(ctypes.byref(s), _ptr(rays), len(rays), int(bool(brute)), _ptr(out), _ptr(boxes), _ptr(prims))
    return out, boxes, prims


def ref_render(scene, camera, width, height, seed, batches, brute=False, debug=0, direct_light=False):
    L = ref()
    s, keep = scene_struct(scene)
    cam = np.ascontiguousarray(np.asarray(camera).reshape(()))
    b = np.asarray(batches, np.int32)
    img = np.zeros((width * height, 3), np.float32)
    rays = ctypes.c_uint64()
    rc = L.ref_render(ctypes.byref(s), _ptr(cam), width, height, int(seed), int(bool(brute)), int(debug),
                      int(bool(direct_light)), _ptr(b), len(b), _ptr(img), ctypes.byref(rays))
    if rc != 0:
        raise RuntimeError(f"reference render failed ({rc})")
    return img, rays.value


def ref_scene(handle, dtypes):
    """Copy a reference-exported scene into numpy arrays of the given dtypes."""
    L = ref()
    counts = (ctypes.c_int64 * 6)()
    rd = (ctypes.c_int32 * 2)()
    L.ref_scene_counts(ctypes.c_void_p(handle), counts, rd)
    arrs = [np.zeros(counts[i], dt) for i, dt in enumerate(dtypes[:6])]
    cam = np.zeros((), dtypes[6])
    L.ref_scene_copy(ctypes.c_void_p(handle), *[_ptr(a) for a in arrs], cam.ctypes.data)
    L.ref_scene_free(ctypes.c_void_p(handle))
    return arrs, int(rd[0]), int(rd[1]), cam
