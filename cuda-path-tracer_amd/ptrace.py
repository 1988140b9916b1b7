"""Python binding of the MI355X path-tracing core (ctypes over include/pt_render.h).

Mirrors the reference renderer's class cuda_renderer and free functions
(src/cuda/cuda_renderer.h:8-32): same method names, argument meaning and
error behaviour -- bad arguments raise ``InvalidArgument`` (a ``ValueError``,
the reference's std::invalid_argument), everything else raises
``RuntimeError`` (std::runtime_error).

Records are numpy structured arrays whose layouts are byte-identical to the
reference's transfer records (src/flat/flat_data.h:14-40).  The native
library must be built (cuda-path-tracer_amd/build.py); importing this module
does not build it and nothing here falls back to a CPU path.
"""
import ctypes
import os
from dataclasses import dataclass, field

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

# ---------------------------------------------------------------- records
VEC3 = ("<f4", (3,))
TRIANGLE = np.dtype([("v0", *VEC3), ("v1", *VEC3), ("v2", *VEC3), ("n0", *VEC3), ("n1", *VEC3), ("n2", *VEC3),
                     ("uv0", *VEC3), ("uv1", *VEC3), ("uv2", *VEC3), ("material_id", "<i4"), ("smooth", "<i4")])
SPHERE = np.dtype([("center", *VEC3), ("radius", "<f4"), ("material_id", "<i4")])
NODE = np.dtype([("minimum", *VEC3), ("maximum", *VEC3), ("type", "<i4"), ("left", "<i4"), ("right", "<i4"),
                 ("primitive", "<i4")])
MATERIAL = np.dtype([("albedo", *VEC3), ("fuzz", "<f4"), ("ref_idx", "<f4"), ("type", "<i4"), ("texture_id", "<i4")])
TEXTURE = np.dtype([("color", *VEC3), ("scale", *VEC3), ("offset", *VEC3), ("type", "<i4"), ("image_id", "<i4"),
                    ("width", "<i4"), ("height", "<i4"), ("byte_offset", "<i4"), ("clamp", "<i4")])
CAMERA = np.dtype([("origin", *VEC3), ("lower_left_corner", *VEC3), ("horizontal", *VEC3), ("vertical", *VEC3),
                   ("u", *VEC3), ("v", *VEC3), ("lens_radius", "<f4"), ("time0", "<f4"), ("time1", "<f4")])
RAY = np.dtype([("origin", *VEC3), ("direction", *VEC3), ("time", "<f4")])
HIT = np.dtype([("t", "<f4"), ("u", "<f4"), ("v", "<f4"), ("p", *VEC3), ("normal", *VEC3), ("material_id", "<i4"),
                ("primitive_type", "<i4"), ("primitive_id", "<i4")])
TRACE_RESULT = np.dtype([("rec", HIT), ("hit", "<i4"), ("error", "<i4")])
for _dt, _size in ((TRIANGLE, 116), (SPHERE, 20), (NODE, 40), (MATERIAL, 28), (TEXTURE, 60), (CAMERA, 84),
                   (RAY, 28), (HIT, 48), (TRACE_RESULT, 56)):
    assert _dt.itemsize == _size, (_dt, _size)

BRANCH, TRIANGLE_PRIM, SPHERE_PRIM = 0, 1, 2
DIFFUSE, METAL, GLASS, LIGHT = 0, 1, 2, 3
CONSTANT, IMAGE, MAP = 0, 1, 2
DEBUG_SHADED, DEBUG_NORMALS, DEBUG_TRIANGLE, DEBUG_MATERIAL, DEBUG_BVH, DEBUG_INTERSECTIONS = range(6)
DEBUG_MODES = {"shaded": 0, "normals": 1, "triangle-id": 2, "material-id": 3, "bvh": 4, "intersections": 5}

PT_OK, PT_EINVAL, PT_ESCENE, PT_EDEVICE, PT_EOVERFLOW, PT_ENOMEM = range(6)


class InvalidArgument(ValueError):
    """The reference's std::invalid_argument."""


@dataclass
class FlatScene:
    """Host arrays of one flat_scene (src/flat/flat_data.h:47-56)."""
    triangles: np.ndarray = field(default_factory=lambda: np.zeros(0, TRIANGLE))
    spheres: np.ndarray = field(default_factory=lambda: np.zeros(0, SPHERE))
    nodes: np.ndarray = field(default_factory=lambda: np.zeros(0, NODE))
    materials: np.ndarray = field(default_factory=lambda: np.zeros(0, MATERIAL))
    textures: np.ndarray = field(default_factory=lambda: np.zeros(0, TEXTURE))
    pixels: np.ndarray = field(default_factory=lambda: np.zeros(0, np.uint8))
    root: int = -1
    max_depth: int = 0

    def normalized(self):
        out = FlatScene(
            np.ascontiguousarray(self.triangles, TRIANGLE), np.ascontiguousarray(self.spheres, SPHERE),
            np.ascontiguousarray(self.nodes, NODE), np.ascontiguousarray(self.materials, MATERIAL),
            np.ascontiguousarray(self.textures, TEXTURE), np.ascontiguousarray(self.pixels, np.uint8),
            int(self.root), int(self.max_depth))
        return out

    def arrays(self):
        return {"triangles": self.triangles, "spheres": self.spheres, "nodes": self.nodes,
                "materials": self.materials, "textures": self.textures, "pixels": self.pixels,
                "root": np.int32(self.root), "max_depth": np.int32(self.max_depth)}


class _SceneDesc(ctypes.Structure):
    _fields_ = [("triangles", ctypes.c_void_p), ("triangle_count", ctypes.c_int64),
                ("spheres", ctypes.c_void_p), ("sphere_count", ctypes.c_int64),
                ("nodes", ctypes.c_void_p), ("node_count", ctypes.c_int64),
                ("materials", ctypes.c_void_p), ("material_count", ctypes.c_int64),
                ("textures", ctypes.c_void_p), ("texture_count", ctypes.c_int64),
                ("pixels", ctypes.c_void_p), ("pixel_bytes", ctypes.c_int64),
                ("root", ctypes.c_int32), ("max_depth", ctypes.c_int32)]


class _Settings(ctypes.Structure):
    _fields_ = [("width", ctypes.c_int32), ("height", ctypes.c_int32), ("seed", ctypes.c_uint64),
                ("brute", ctypes.c_int32), ("debug", ctypes.c_int32), ("direct_light", ctypes.c_int32),
                ("max_path_depth", ctypes.c_int32)]


class LaunchStats(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int32), ("trace_launches", ctypes.c_int32), ("trace_ms", ctypes.c_float),
                ("shade_ms", ctypes.c_float), ("extend_rays", ctypes.c_uint64), ("shadow_rays", ctypes.c_uint64),
                ("regen_ms", ctypes.c_float),
                ("box_tests", ctypes.c_uint64), ("triangle_tests", ctypes.c_uint64), ("sphere_tests", ctypes.c_uint64),
                ("drained_paths", ctypes.c_uint64), ("drain_ms", ctypes.c_float), ("split_launches", ctypes.c_int32),
                ("retraced_queries", ctypes.c_uint64), ("hit_queries", ctypes.c_uint64),
                ("retrace_box_tests", ctypes.c_uint64), ("retrace_primitive_tests", ctypes.c_uint64)]


class Tuning(ctypes.Structure):
    _fields_ = [("trace_block", ctypes.c_int32), ("shade_block", ctypes.c_int32), ("blocks_per_cu", ctypes.c_int32),
                ("check_interval", ctypes.c_int32), ("time_kernels", ctypes.c_int32), ("count_tests", ctypes.c_int32),
                ("exact_traversal", ctypes.c_int32), ("drain_threshold", ctypes.c_int32), ("lane_refill", ctypes.c_int32),
                ("partitions", ctypes.c_int32), ("handoff_steps", ctypes.c_int32),
                ("path_shade_batch", ctypes.c_int32), ("path_blocks_per_cu", ctypes.c_int32),
                ("path_first_samples", ctypes.c_int32), ("path_sort_group", ctypes.c_int32),
                ("path_trace_steps", ctypes.c_int32)]


class BatchTrial(ctypes.Structure):
    _fields_ = [("tuning", Tuning), ("mrays", ctypes.c_float)]


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data) if a.size else ctypes.c_void_p(0)


def scene_desc(scene):
    """ctypes pt_scene_desc viewing ``scene``'s arrays (keep ``scene`` alive)."""
    return _SceneDesc(_ptr(scene.triangles), len(scene.triangles), _ptr(scene.spheres), len(scene.spheres),
                      _ptr(scene.nodes), len(scene.nodes), _ptr(scene.materials), len(scene.materials),
                      _ptr(scene.textures), len(scene.textures), _ptr(scene.pixels), len(scene.pixels),
                      int(scene.root), int(scene.max_depth))


# ---------------------------------------------------------------- libraries
_LIB = None
_SCENE_LIB = None
LIBRARY_PATH = os.path.join(HERE, "libptrace.so")
SCENE_LIBRARY_PATH = os.path.join(HERE, "libptscene.so")

PT_FUNCTIONS = {
    "pt_last_error": (ctypes.c_char_p, []),
    "pt_device_available": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_size_t]),
    "pt_host_alloc": (ctypes.c_int, [ctypes.c_size_t, ctypes.c_void_p]),
    "pt_host_free": (ctypes.c_int, [ctypes.c_void_p]),
    "pt_build_bvh": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p,
                                    ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "pt_renderer_create": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "pt_renderer_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "pt_renderer_reset": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "pt_renderer_add_samples": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32]),
    "pt_renderer_add_samples_noalloc": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32]),
    "pt_renderer_set_block_shape": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32]),
    "pt_renderer_read_image": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t]),
    "pt_renderer_read_rgba": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t]),
    "pt_renderer_write_rgba_device": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t]),
    "pt_renderer_stats": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "pt_query_rays": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int32, ctypes.c_void_p]),
    "pt_query_rays_stats": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int32,
                                           ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "pt_query_rays_wide": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p,
                                          ctypes.c_void_p]),
    "pt_renderer_read_state": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t]),
    "pt_renderer_write_state": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
                                               ctypes.c_int32, ctypes.c_uint64]),
    "pt_renderer_last_launch_stats": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "pt_renderer_set_tuning": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "pt_renderer_iteration_log": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
                                                 ctypes.c_void_p]),
    "pt_renderer_enable_aovs": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32]),
    "pt_batch_create": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32,
                                       ctypes.c_int32, ctypes.c_void_p]),
    "pt_batch_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "pt_batch_reset": (ctypes.c_int, [ctypes.c_void_p]),
    "pt_batch_add_samples": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32]),
    "pt_batch_read_image": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p, ctypes.c_size_t]),
    "pt_batch_stats": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "pt_batch_last_launch_stats": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "pt_batch_get_tuning": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "pt_batch_set_tuning": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "pt_batch_autotune": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, ctypes.c_void_p,
                                         ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p]),
    "pt_renderer_read_aovs": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t]),
}


def lib():
    """The loaded libptrace.so; raises if it has not been built."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIBRARY_PATH):
            raise RuntimeError(f"HIP library not built: {LIBRARY_PATH} (run cuda-path-tracer_amd/build.py)")
        handle = ctypes.CDLL(LIBRARY_PATH)
        for name, (res, args) in PT_FUNCTIONS.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = handle
    return _LIB


def scene_lib():
    global _SCENE_LIB
    if _SCENE_LIB is None:
        if not os.path.exists(SCENE_LIBRARY_PATH):
            raise RuntimeError(f"scene library not built: {SCENE_LIBRARY_PATH} (run cuda-path-tracer_amd/build.py)")
        h = ctypes.CDLL(SCENE_LIBRARY_PATH)
        for name in ("pts_cornell", "pts_icosphere", "pts_soup", "pts_build"):
            getattr(h, name).restype = ctypes.c_void_p
        h.pts_cornell.argtypes = [ctypes.c_uint64, ctypes.c_int32, ctypes.c_int32, ctypes.c_void_p]
        h.pts_icosphere.argtypes = [ctypes.c_int32, ctypes.c_uint64, ctypes.c_int32, ctypes.c_int32, ctypes.c_void_p]
        h.pts_soup.argtypes = [ctypes.c_int64, ctypes.c_uint64, ctypes.c_int32, ctypes.c_int32, ctypes.c_void_p]
        h.pts_build.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p,
                                ctypes.c_int64, ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_int64,
                                ctypes.c_int32]
        h.pts_counts.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        h.pts_copy.argtypes = [ctypes.c_void_p] * 7
        h.pts_free.argtypes = [ctypes.c_void_p]
        h.pts_camera.argtypes = [ctypes.c_void_p] * 3 + [ctypes.c_float] * 6 + [ctypes.c_void_p]
        h.pt_scene_last_error.restype = ctypes.c_char_p
        h.pts_build_seeded.restype = ctypes.c_void_p
        h.pts_build_seeded.argtypes = h.pts_build.argtypes + [ctypes.c_int64]
        h.pts_load_obj.restype = ctypes.c_void_p
        h.pts_load_obj.argtypes = [ctypes.c_char_p, ctypes.c_void_p, ctypes.c_void_p]
        h.pts_warnings.restype = ctypes.c_char_p
        h.pts_warnings.argtypes = [ctypes.c_void_p]
        h.pts_load_image.argtypes = [ctypes.c_char_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        h.pts_load_image.restype = ctypes.c_int32
        h.pts_free_image.argtypes = [ctypes.c_void_p]
        h.pts_write_ppm.argtypes = [ctypes.c_char_p, ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32]
        h.pts_write_ppm.restype = ctypes.c_int32
        _SCENE_LIB = h
    return _SCENE_LIB


def _check(status):
    if status == PT_OK:
        return
    message = lib().pt_last_error().decode()
    if status == PT_EINVAL:
        raise InvalidArgument(message)
    raise RuntimeError(message)


def device_available():
    """cuda_device_available (cuda_renderer.cu:128-140): (available, description)."""
    buf = ctypes.create_string_buffer(512)
    ok = lib().pt_device_available(buf, 512) == PT_OK
    return ok, buf.value.decode()


def pinned(shape, dtype=np.float32):
    """A zeroed numpy array in page-locked host memory (pt_host_alloc), freed
    with the array: device-to-host reads into it (Renderer.image(out=...),
    the denoiser's output) run at the link's rate."""
    dtype = np.dtype(dtype)
    count = int(np.prod(shape))
    nbytes = max(1, count * dtype.itemsize)
    ptr = ctypes.c_void_p()
    _check(lib().pt_host_alloc(nbytes, ctypes.byref(ptr)))
    buf = (ctypes.c_uint8 * nbytes).from_address(ptr.value)
    arr = np.frombuffer(buf, dtype=dtype, count=count).reshape(shape)
    arr[...] = 0
    import weakref
    weakref.finalize(buf, lib().pt_host_free, ctypes.c_void_p(ptr.value))
    return arr


# ---------------------------------------------------------------- renderer
class Renderer:
    """cuda_renderer (src/cuda/cuda_renderer.h:10-29) over the HIP core.

    ``max_depth`` lifts the reference's fixed path-length limit of 50
    (transport.h:174,193); 0 keeps 50.
    """

    def __init__(self, scene, camera, width, height, seed, brute=False, debug=DEBUG_SHADED, direct_light=False,
                 max_depth=0):
        self._scene = scene.normalized()
        self._camera = np.ascontiguousarray(np.asarray(camera, CAMERA).reshape(()))
        self.width, self.height = int(width), int(height)
        settings = _Settings(self.width, self.height, int(seed) & (2**64 - 1), int(bool(brute)), int(debug),
                             int(bool(direct_light)), int(max_depth))
        desc = scene_desc(self._scene)
        handle = ctypes.c_void_p()
        _check(lib().pt_renderer_create(ctypes.byref(desc), self._camera.ctypes.data, ctypes.byref(settings),
                                        ctypes.byref(handle)))
        self._h = handle

    def close(self):
        if getattr(self, "_h", None):
            lib().pt_renderer_destroy(self._h)
            self._h = None

    __del__ = close

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def pixels(self):
        return self.width * self.height

    def reset(self, camera):
        cam = np.ascontiguousarray(np.asarray(camera, CAMERA).reshape(()))
        _check(lib().pt_renderer_reset(self._h, cam.ctypes.data))

    def add_samples(self, count):
        _check(lib().pt_renderer_add_samples(self._h, int(count)))

    def add_samples_noalloc(self, count):
        """pt_renderer_add_samples_noalloc: add_samples with its scratch in a fixed device array."""
        _check(lib().pt_renderer_add_samples_noalloc(self._h, int(count)))

    def set_block_shape(self, width, height):
        _check(lib().pt_renderer_set_block_shape(self._h, int(width), int(height)))

    def image(self, out=None):
        """Averaged linear RGB, (width*height, 3) float32, row 0 at the bottom;
        into `out` when given (e.g. a pinned() array)."""
        if out is None:
            out = np.empty((self.pixels, 3), np.float32)
        elif out.dtype != np.float32 or out.size < 3 * self.pixels or not out.flags.c_contiguous:
            raise ValueError("out must be a C-contiguous float32 array of at least 3 * pixels values")
        _check(lib().pt_renderer_read_image(self._h, out.ctypes.data, out.size))
        return out

    def rgba_image(self):
        out = np.empty(self.pixels * 4, np.uint8)
        _check(lib().pt_renderer_read_rgba(self._h, out.ctypes.data, out.size))
        return out

    def write_rgba(self, device_pointer, nbytes):
        _check(lib().pt_renderer_write_rgba_device(self._h, ctypes.c_void_p(device_pointer), int(nbytes)))

    def _stats(self):
        s, ms, r = ctypes.c_int32(), ctypes.c_float(), ctypes.c_uint64()
        _check(lib().pt_renderer_stats(self._h, ctypes.byref(s), ctypes.byref(ms), ctypes.byref(r)))
        return s.value, ms.value, r.value

    def sample_count(self):
        return self._stats()[0]

    def kernel_milliseconds(self):
        return self._stats()[1]

    def rays_traced(self):
        return self._stats()[2]

    def read_state(self):
        """(sum (N,3) float32, rng state (N,) uint64): the resumable state."""
        s = np.empty((self.pixels, 3), np.float32)
        r = np.empty(self.pixels, np.uint64)
        _check(lib().pt_renderer_read_state(self._h, s.ctypes.data, r.ctypes.data, self.pixels))
        return s, r

    def write_state(self, sums, rng, samples, rays):
        s = np.ascontiguousarray(sums, np.float32)
        r = np.ascontiguousarray(rng, np.uint64)
        _check(lib().pt_renderer_write_state(self._h, s.ctypes.data, r.ctypes.data, self.pixels, int(samples), int(rays)))

    def last_launch_stats(self):
        st = LaunchStats()
        _check(lib().pt_renderer_last_launch_stats(self._h, ctypes.byref(st)))
        return {k: getattr(st, k) for k, _ in LaunchStats._fields_}

    def enable_aovs(self, enable=True):
        """Accumulate first-hit albedo and shading normal (denoiser features)."""
        _check(lib().pt_renderer_enable_aovs(self._h, int(bool(enable))))

    def aovs(self):
        """(albedo, normal) averaged over the samples, each (W*H, 3) float32."""
        albedo = np.zeros((self.pixels, 3), np.float32)
        normal = np.zeros((self.pixels, 3), np.float32)
        _check(lib().pt_renderer_read_aovs(self._h, albedo.ctypes.data, normal.ctypes.data, albedo.size))
        return albedo, normal

    def iteration_log(self):
        """(extension, shadow) query counts per bounce iteration of the last add_samples."""
        cap = 1 << 16
        e = np.zeros(cap, np.uint32)
        sh = np.zeros(cap, np.uint32)
        n = ctypes.c_int32()
        _check(lib().pt_renderer_iteration_log(self._h, e.ctypes.data, sh.ctypes.data, cap, ctypes.byref(n)))
        m = min(n.value, cap)
        return e[:m], sh[:m]

    def set_tuning(self, trace_block=0, shade_block=0, blocks_per_cu=0, check_interval=0, time_kernels=False,
                   count_tests=False, exact_traversal=False, drain_threshold=0, lane_refill=0, partitions=0,
                   handoff_steps=0, path_shade_batch=0, path_blocks_per_cu=0, path_first_samples=0, path_sort_group=0,
                   path_trace_steps=0):
        t = Tuning(trace_block, shade_block, blocks_per_cu, check_interval, int(bool(time_kernels)),
                   int(bool(count_tests)), int(bool(exact_traversal)), int(drain_threshold), int(lane_refill),
                   int(partitions), int(handoff_steps), int(path_shade_batch), int(path_blocks_per_cu),
                   int(path_first_samples), int(path_sort_group), int(path_trace_steps))
        _check(lib().pt_renderer_set_tuning(self._h, ctypes.byref(t)))



BATCH_VIEW = np.dtype([("seed", "<u8"), ("scene", "<i4"), ("width", "<i4"), ("height", "<i4"), ("camera", CAMERA)])
assert BATCH_VIEW.itemsize == 104


class Batch:
    """Batched rendering of many views of many scenes in one graph-captured
    bounce loop (pt_batch_* in include/pt_render.h).

    scenes: list of FlatScene; views: list of (scene index, camera, width,
    height, seed).  Each view's image equals a single Renderer's."""

    def __init__(self, scenes, views, direct_light=True, max_depth=0):
        self._scenes = [s.normalized() for s in scenes]
        descs = (_SceneDesc * len(scenes))(*[scene_desc(s) for s in self._scenes])
        v = np.zeros(len(views), BATCH_VIEW)
        for i, (scene, cam, w, h, seed) in enumerate(views):
            v[i]["scene"], v[i]["width"], v[i]["height"], v[i]["seed"] = scene, w, h, seed
            v[i]["camera"] = np.asarray(cam, CAMERA)
        self.views = v
        h = ctypes.c_void_p()
        _check(lib().pt_batch_create(descs, len(scenes), v.ctypes.data, len(v), int(bool(direct_light)), int(max_depth),
                                     ctypes.byref(h)))
        self._h = h

    def __del__(self):
        if getattr(self, "_h", None):
            lib().pt_batch_destroy(self._h)
            self._h = None

    def reset(self):
        _check(lib().pt_batch_reset(self._h))

    def add_samples(self, count):
        _check(lib().pt_batch_add_samples(self._h, int(count)))

    def image(self, view):
        w, h = int(self.views[view]["width"]), int(self.views[view]["height"])
        out = np.zeros((w * h, 3), np.float32)
        _check(lib().pt_batch_read_image(self._h, int(view), out.ctypes.data, out.size))
        return out

    def stats(self):
        n, ms, rays = ctypes.c_int32(), ctypes.c_float(), ctypes.c_uint64()
        _check(lib().pt_batch_stats(self._h, ctypes.byref(n), ctypes.byref(ms), ctypes.byref(rays)))
        return {"samples": n.value, "milliseconds": ms.value, "rays": rays.value}

    def last_launch_stats(self):
        st = LaunchStats()
        _check(lib().pt_batch_last_launch_stats(self._h, ctypes.byref(st)))
        return {k: getattr(st, k) for k, _ in LaunchStats._fields_}

    def tuning(self):
        t = Tuning()
        _check(lib().pt_batch_get_tuning(self._h, ctypes.byref(t)))
        return {k: getattr(t, k) for k, _ in Tuning._fields_}

    def set_tuning(self, **fields):
        t = Tuning(**{k: int(v) for k, v in fields.items()})
        _check(lib().pt_batch_set_tuning(self._h, ctypes.byref(t)))

    def autotune(self, probe_samples=1, passes=1, capacity=256):
        """Coordinate-descent autotuner; returns (best tuning, best Mrays/s, trials)."""
        best, rate, count = Tuning(), ctypes.c_float(), ctypes.c_int32()
        trials = (BatchTrial * capacity)()
        _check(lib().pt_batch_autotune(self._h, int(probe_samples), int(passes), ctypes.byref(best), ctypes.byref(rate),
                                       trials, capacity, ctypes.byref(count)))
        rows = [dict({k: getattr(t.tuning, k) for k, _ in Tuning._fields_}, mrays=t.mrays)
                for t in trials[:min(count.value, capacity)]]
        return {k: getattr(best, k) for k, _ in Tuning._fields_}, rate.value, rows


def query_rays(scene, rays, brute=False, stats=False):
    """cuda_query_rays (cuda_renderer.cu:248-260); with ``stats`` also the
    per-ray box and primitive test counts (trace_stats)."""
    scene = scene.normalized()
    rays = np.ascontiguousarray(rays, RAY).reshape(-1)
    out = np.zeros(len(rays), TRACE_RESULT)
    desc = scene_desc(scene)
    if stats:
        boxes = np.zeros(len(rays), np.int32)
        prims = np.zeros(len(rays), np.int32)
        _check(lib().pt_query_rays_stats(ctypes.byref(desc), _ptr(rays), len(rays), int(bool(brute)), _ptr(out),
                                         _ptr(boxes), _ptr(prims)))
        return out, boxes, prims
    _check(lib().pt_query_rays(ctypes.byref(desc), _ptr(rays), len(rays), int(bool(brute)), _ptr(out)))
    return out


def query_rays_wide(scene, rays):
    """(results, retraced): the closest hits as the shaded path's 4-wide
    traversal finds them (equal to query_rays'), and where its
    reference-order retrace ran."""
    scene = scene.normalized()
    rays = np.ascontiguousarray(rays, RAY).reshape(-1)
    out = np.zeros(len(rays), TRACE_RESULT)
    flags = np.zeros(len(rays), np.int32)
    desc = scene_desc(scene)
    _check(lib().pt_query_rays_wide(ctypes.byref(desc), _ptr(rays), len(rays), _ptr(out), _ptr(flags)))
    return out, flags.astype(bool)


# ---------------------------------------------------------------- scenes
def _take_scene(handle, warnings=None):
    L = scene_lib()
    if not handle:
        raise InvalidArgument(L.pt_scene_last_error().decode())
    if warnings is not None:
        warnings.append(L.pts_warnings(ctypes.c_void_p(handle)).decode())
    counts = (ctypes.c_int64 * 6)()
    rd = (ctypes.c_int32 * 2)()
    L.pts_counts(ctypes.c_void_p(handle), counts, rd)
    s = FlatScene(np.zeros(counts[0], TRIANGLE), np.zeros(counts[1], SPHERE), np.zeros(counts[2], NODE),
                  np.zeros(counts[3], MATERIAL), np.zeros(counts[4], TEXTURE), np.zeros(counts[5], np.uint8),
                  rd[0], rd[1])
    L.pts_copy(ctypes.c_void_p(handle), _ptr(s.triangles), _ptr(s.spheres), _ptr(s.nodes), _ptr(s.materials),
               _ptr(s.textures), _ptr(s.pixels))
    L.pts_free(ctypes.c_void_p(handle))
    return s


def cornell(seed=7, width=320, height=200):
    """Cornell-type room with three seeded spheres; returns (scene, camera)."""
    cam = np.zeros((), CAMERA)
    h = scene_lib().pts_cornell(int(seed), int(width), int(height), cam.ctypes.data)
    return _take_scene(h), cam


def icosphere(frequency=224, seed=7, width=1920, height=1080):
    """Displaced geodesic icosphere of 20*frequency^2 triangles; (scene, camera)."""
    cam = np.zeros((), CAMERA)
    h = scene_lib().pts_icosphere(int(frequency), int(seed), int(width), int(height), cam.ctypes.data)
    return _take_scene(h), cam


def soup(triangles, seed=7, width=640, height=480):
    cam = np.zeros((), CAMERA)
    h = scene_lib().pts_soup(int(triangles), int(seed), int(width), int(height), cam.ctypes.data)
    return _take_scene(h), cam


SPLITS = {"sah": 0, "median": 1, "reference": 2}

OBJ_OPTIONS = np.dtype([("scale", "<f4", 3), ("rotate", "<f4", 3), ("translate", "<f4", 3), ("aspect", "<f4"),
                        ("split", "<i4"), ("seed", "<i8")], align=True)
OBJ_INFO = np.dtype([("lookfrom", "<f4", 3), ("lookat", "<f4", 3), ("focus_dist", "<f4"),
                     ("source_vertices", "<i8"), ("shape_count", "<i8"), ("mesh_triangles", "<i8")], align=True)


def build_scene(triangles=None, spheres=None, materials=None, textures=None, pixels=None, split="sah", seed=0):
    """BVH over the given primitives in the flat format (leaf-ordered).

    split "reference" rebuilds the reference's own hierarchy (random-axis
    median split driven by drand48 after srand48(seed), src/bvh/bvh.h:132-171)
    and numbers materials and textures as its exporter does."""
    t = np.ascontiguousarray(triangles if triangles is not None else np.zeros(0, TRIANGLE), TRIANGLE)
    s = np.ascontiguousarray(spheres if spheres is not None else np.zeros(0, SPHERE), SPHERE)
    m = np.ascontiguousarray(materials if materials is not None else np.zeros(0, MATERIAL), MATERIAL)
    x = np.ascontiguousarray(textures if textures is not None else np.zeros(0, TEXTURE), TEXTURE)
    p = np.ascontiguousarray(pixels if pixels is not None else np.zeros(0, np.uint8), np.uint8)
    h = scene_lib().pts_build_seeded(_ptr(t), len(t), _ptr(s), len(s), _ptr(m), len(m), _ptr(x), len(x), _ptr(p),
                                     len(p), SPLITS[split], int(seed))
    return _take_scene(h)


def build_bvh_device(triangles=None, spheres=None, materials=None, textures=None, pixels=None):
    """Linear BVH built on the device (pt_build_bvh): a FlatScene over the
    given arrays, primitives in their given order, and the device time of
    the build in milliseconds."""
    t = np.ascontiguousarray(triangles if triangles is not None else np.zeros(0, TRIANGLE), TRIANGLE)
    s = np.ascontiguousarray(spheres if spheres is not None else np.zeros(0, SPHERE), SPHERE)
    n = len(t) + len(s)
    nodes = np.zeros(max(2 * n - 1, 0), NODE)
    root, depth, ms = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_float()
    _check(lib().pt_build_bvh(_ptr(t), len(t), _ptr(s), len(s), _ptr(nodes), len(nodes), ctypes.byref(root),
                              ctypes.byref(depth), ctypes.byref(ms)))
    scene = FlatScene(t, s, nodes,
                      np.ascontiguousarray(materials if materials is not None else np.zeros(0, MATERIAL), MATERIAL),
                      np.ascontiguousarray(textures if textures is not None else np.zeros(0, TEXTURE), TEXTURE),
                      np.ascontiguousarray(pixels if pixels is not None else np.zeros(0, np.uint8), np.uint8),
                      root.value, depth.value)
    return scene, ms.value


def load_obj(path, aspect, scale=(1, 1, 1), rotate=(0, 0, 0), translate=(0, 0, 0), split="reference", seed=7,
             vfov=40.0, aperture=0.0, time0=0.0, time1=1.0):
    """Import an OBJ/MTL asset staged as the reference's obj_scene
    (src/scene/obj_scene.h:10-31): returns (scene, camera, info) with info a
    dict of lookfrom, lookat, focus_dist, source_vertices, shape_count,
    mesh_triangles and the loader's warnings.  With split "reference" the
    arrays equal the reference's srand48(seed); obj_scene; flatten_scene."""
    opt = np.zeros((), OBJ_OPTIONS)
    opt["scale"], opt["rotate"], opt["translate"] = scale, rotate, translate
    opt["aspect"], opt["split"], opt["seed"] = aspect, SPLITS[split], int(seed)
    info = np.zeros((), OBJ_INFO)
    h = scene_lib().pts_load_obj(os.fsencode(path), opt.ctypes.data, info.ctypes.data)
    warnings = []
    scene = _take_scene(h, warnings)
    camera = make_camera(info["lookfrom"], info["lookat"], (0, 1, 0), vfov, aspect, aperture, float(info["focus_dist"]),
                         time0, time1)
    out = {k: (info[k].copy() if info[k].ndim else info[k].item()) for k in OBJ_INFO.names}
    out["warnings"] = warnings[0]
    return scene, camera, out


def load_image(path):
    """Decode a PNG, binary PNM or JPEG texture to an (height, width, 3) uint8 array."""
    L = scene_lib()
    w, h, buf = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_void_p()
    if L.pts_load_image(os.fsencode(path), ctypes.byref(w), ctypes.byref(h), ctypes.byref(buf)) != 0:
        raise RuntimeError(L.pt_scene_last_error().decode())
    try:
        data = np.ctypeslib.as_array(ctypes.cast(buf, ctypes.POINTER(ctypes.c_uint8)), (h.value, w.value, 3)).copy()
    finally:
        L.pts_free_image(buf)
    return data


def write_ppm(path, image, width, height):
    """Write a linear RGB image (row 0 at the bottom) as the reference
    application's ASCII PPM; path "-" is stdout."""
    rgb = np.ascontiguousarray(image, np.float32).reshape(-1)
    if rgb.size != 3 * width * height:
        raise InvalidArgument("image size does not match width*height")
    L = scene_lib()
    if L.pts_write_ppm(os.fsencode(path), rgb.ctypes.data, int(width), int(height)) != 0:
        raise RuntimeError(L.pt_scene_last_error().decode())


def make_camera(lookfrom, lookat, vup, vfov, aspect, aperture, focus_dist, time0=0.0, time1=1.0):
    """Camera basis of the reference's camera constructor (src/core/camera.h:16-43)."""
    f = np.asarray(lookfrom, np.float32)
    a = np.asarray(lookat, np.float32)
    u = np.asarray(vup, np.float32)
    cam = np.zeros((), CAMERA)
    scene_lib().pts_camera(f.ctypes.data, a.ctypes.data, u.ctypes.data, vfov, aspect, aperture, focus_dist, time0,
                           time1, cam.ctypes.data)
    return cam


def save_scene(path, scene, camera=None, **extra):
    arrays = scene.arrays()
    if camera is not None:
        arrays["camera"] = np.asarray(camera, CAMERA)
    arrays.update(extra)
    np.savez_compressed(path, **arrays)


def load_scene(path):
    """(scene, camera-or-None, extra arrays) from save_scene's npz."""
    with np.load(path, allow_pickle=False) as z:
        scene = FlatScene(z["triangles"], z["spheres"], z["nodes"], z["materials"], z["textures"], z["pixels"],
                          int(z["root"]), int(z["max_depth"]))
        camera = z["camera"] if "camera" in z.files else None
        extra = {k: z[k] for k in z.files if k not in scene.arrays() and k != "camera"}
    return scene, camera, extra
