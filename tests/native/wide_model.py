"""ctypes access to tests/native/libwidemodel.so, the CPU model of the shaded
path's traversal (test infrastructure): the renderer's 4-wide traversal, the
same without its reference-order fallback, and the reference-order traversal,
run on the host over the layouts the renderer uploads."""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
WIDE, WIDE_NO_FALLBACK, REFERENCE_ORDER = 0, 1, 2
_LIBS = {}


def lib(sanitized=False):
    name = "libwidemodel_asan.so" if sanitized else "libwidemodel.so"
    if name not in _LIBS:
        path = os.path.join(HERE, name)
        if not os.path.exists(path):
            import build_model
            build_model.build(sanitize=sanitized)
        so = ctypes.CDLL(path)
        so.wm_layout_info.restype = ctypes.c_int
        so.wm_layout_info.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_char_p, ctypes.c_size_t]
        so.wm_trace.restype = ctypes.c_int64
        so.wm_trace.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32,
                                ctypes.c_int32, ctypes.c_void_p, ctypes.c_void_p]
        _LIBS[name] = so
    return _LIBS[name]


def pow5(x, sanitized=False):
    """(correctly rounded x^5, the renderer's host-libm x^5, table size) for float32 x."""
    x = np.ascontiguousarray(x, np.float32)
    exact = np.zeros(len(x), np.float64)
    libm = np.zeros(len(x), np.float64)
    so = lib(sanitized)
    so.wm_pow5.restype = ctypes.c_int64
    so.wm_pow5.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p]
    n = so.wm_pow5(x.ctypes.data, len(x), exact.ctypes.data, libm.ctypes.data)
    return exact, libm, int(n)


def layout_info(scene, sanitized=False):
    """dict: wide depth, wide stack capacity, binary stack capacity, wide nodes, leaves, has_wide."""
    import ptrace
    scene = scene.normalized()
    desc = ptrace.scene_desc(scene)
    info = np.zeros(6, np.int32)
    err = ctypes.create_string_buffer(256)
    if lib(sanitized).wm_layout_info(ctypes.byref(desc), info.ctypes.data, err, 256) != 0:
        raise RuntimeError(err.value.decode())
    keys = ("wide_depth", "wide_capacity", "binary_capacity", "wide_nodes", "leaves", "has_wide")
    return dict(zip(keys, (int(v) for v in info)))


def trace(scene, rays, mode=WIDE, any_hit=False, high=None, sanitized=False):
    """(hits, stats, overflows): hits[:, 0..2] t/u/v bits, hits[:, 3] code
    (primitive*2 + sphere, -1 for none); stats[:, 0..3] nodes visited,
    primitives tested, deepest stack entry + 1, retraced."""
    import ptrace
    scene = scene.normalized()
    desc = ptrace.scene_desc(scene)
    rays = np.ascontiguousarray(rays, ptrace.RAY)
    n = len(rays)
    hi = np.full(n, np.finfo(np.float32).max, np.float32) if high is None else np.ascontiguousarray(high, np.float32)
    hits = np.zeros((n, 4), np.int32)
    stats = np.zeros((n, 4), np.int32)
    ov = lib(sanitized).wm_trace(ctypes.byref(desc), rays.ctypes.data, hi.ctypes.data, n, mode, int(any_hit),
                                 hits.ctypes.data, stats.ctypes.data)
    if ov < 0:
        raise RuntimeError("invalid flat scene data")
    return hits, stats, int(ov)
