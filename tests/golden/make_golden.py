"""Write the golden fixtures under tests/golden/ from the compiled reference.

Run in the container that holds the reference tree, after
``bash oracle/build_ref.sh``:  ``python tests/golden/make_golden.py``.

Every expected output here comes from the reference's own host path
(transport.h via portable_renderer / trace_world, compiled from the sources
under /root/reference by oracle/build_ref.sh).  Scenes are either the
reference's own exports (its OBJ loader, random-axis BVH and flat exporter,
seeded with srand48 as in the reference tests) or scenes built by this
project's builder; both are stored as flat arrays, so the fixtures do not
depend on the reference tree at test time.
"""
import ctypes
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(ROOT, "cuda-path-tracer_amd"), os.path.join(ROOT, "oracle")]

import oracle  # noqa: E402
import ptrace  # noqa: E402

REF_ROOT = os.environ.get("PT_REFERENCE_ROOT", "/root/reference")
DTYPES = [ptrace.TRIANGLE, ptrace.SPHERE, ptrace.NODE, ptrace.MATERIAL, ptrace.TEXTURE, np.uint8, ptrace.CAMERA]
MODES = range(6)


def ref_scene(handle):
    arrs, root, depth, cam = oracle.ref_scene(handle, DTYPES)
    return ptrace.FlatScene(*arrs, root, depth), cam


def load_asset(name, vfov=40.0, aperture=0.0, t0=0.0, t1=1.0, seed=7,
               scale=(1, 1, 1), rotate=(0, 0, 0), translate=(0, 0, 0)):
    L = oracle.ref()
    path = os.path.join(REF_ROOT, "assets", "models", name).encode()
    f = lambda v: np.asarray(v, np.float32)  # noqa: E731
    s, r, t = f(scale), f(rotate), f(translate)
    h = L.ref_load_obj(path, 1.5, seed, s.ctypes.data, r.ctypes.data, t.ctypes.data, vfov, aperture, t0, t1)
    if not h:
        raise RuntimeError(f"reference could not load {name}")
    return ref_scene(h)


def query_rays(scene, camera, n, seed):
    """Camera rays (half), random rays (quarter) and axis-aligned rays (quarter)."""
    rng = np.random.default_rng(seed)
    rays = np.zeros(n, ptrace.RAY)
    lo = np.array([scene.nodes[scene.root]["minimum"]], np.float32)[0] if len(scene.nodes) else np.zeros(3, np.float32)
    hi = np.array([scene.nodes[scene.root]["maximum"]], np.float32)[0] if len(scene.nodes) else np.ones(3, np.float32)
    centre, extent = (lo + hi) / 2, np.minimum(hi - lo, 50.0)
    cam = np.asarray(camera)
    k = n // 2
    u = rng.random(k).astype(np.float32)[:, None]
    v = rng.random(k).astype(np.float32)[:, None]
    rays["origin"][:k] = cam["origin"]
    rays["direction"][:k] = cam["lower_left_corner"] + u * cam["horizontal"] + v * cam["vertical"] - cam["origin"]
    m = (n - k) // 2
    rays["origin"][k:k + m] = centre + (rng.random((m, 3)) - 0.5).astype(np.float32) * extent
    rays["direction"][k:k + m] = rng.normal(size=(m, 3)).astype(np.float32)
    rest = n - k - m
    rays["origin"][k + m:] = centre + (rng.random((rest, 3)) - 0.5).astype(np.float32) * extent
    d = np.zeros((rest, 3), np.float32)
    d[np.arange(rest), np.arange(rest) % 3] = np.where(np.arange(rest) % 2, 1.0, -1.0)
    rays["direction"][k + m:] = d
    rays["time"] = 0.42
    return rays


def render_all(scene, camera, w, h, seed, batches, modes=MODES, direct=(False, True)):
    out = {}
    for mode in modes:
        for dl in (direct if mode == 0 else (False,)):
            img, rays = oracle.ref_render(scene, camera, w, h, seed, batches, False, mode, dl)
            out[f"image_m{mode}_d{int(dl)}"] = img
            out[f"rays_m{mode}_d{int(dl)}"] = np.uint64(rays)
    return out


def save(name, scene, camera, **extra):
    path = os.path.join(HERE, name)
    ptrace.save_scene(path, scene, camera, **extra)
    print("wrote", path, os.path.getsize(path), "bytes")


ASSETS = os.path.join(HERE, "assets", "models")


def png_chunk(kind, body):
    import struct
    import zlib
    return struct.pack(">I", len(body)) + kind + body + struct.pack(">I", zlib.crc32(kind + body) & 0xffffffff)


def encode_png(samples, color, depth, interlace=False, level=6, strategy=None, palette=None):
    """Minimal PNG writer for decoder fixtures: samples is (h, w, channels)
    of integers below 2^depth; rows are filtered with filter type row % 5."""
    import struct
    import zlib
    h, w, c = samples.shape

    def pack_rows(img):
        rows = []
        for y in range(img.shape[0]):
            vals = img[y].reshape(-1).astype(np.uint32)
            if depth == 16:
                raw = b"".join(struct.pack(">H", int(v)) for v in vals)
            elif depth == 8:
                raw = bytes(int(v) for v in vals)
            else:
                bits = "".join(format(int(v), f"0{depth}b") for v in vals)
                bits += "0" * (-len(bits) % 8)
                raw = bytes(int(bits[i:i + 8], 2) for i in range(0, len(bits), 8))
            rows.append(raw)
        return rows

    def filter_rows(rows):
        bpp = max(1, c * depth // 8)
        out = b""
        prev = bytes(len(rows[0])) if rows else b""
        for y, raw in enumerate(rows):
            f = y % 5
            line = bytearray()
            for i, x in enumerate(raw):
                a = raw[i - bpp] if i >= bpp else 0
                b = prev[i]
                cc = prev[i - bpp] if i >= bpp else 0
                if f == 0:
                    p = 0
                elif f == 1:
                    p = a
                elif f == 2:
                    p = b
                elif f == 3:
                    p = (a + b) >> 1
                else:
                    pp = a + b - cc
                    pa, pb, pc = abs(pp - a), abs(pp - b), abs(pp - cc)
                    p = a if (pa <= pb and pa <= pc) else (b if pb <= pc else cc)
                line.append((x - p) & 255)
            out += bytes([f]) + bytes(line)
            prev = raw
        return out

    if interlace:
        data = b""
        for x0, y0, dx, dy in ((0, 0, 8, 8), (4, 0, 8, 8), (0, 4, 4, 8), (2, 0, 4, 4), (0, 2, 2, 4), (1, 0, 2, 2),
                               (0, 1, 1, 2)):
            sub = samples[y0::dy, x0::dx]
            if sub.shape[0] and sub.shape[1]:
                data += filter_rows(pack_rows(sub))
    else:
        data = filter_rows(pack_rows(samples))
    comp = zlib.compressobj(level, zlib.DEFLATED, 15, 9, strategy if strategy is not None else zlib.Z_DEFAULT_STRATEGY)
    z = comp.compress(data) + comp.flush()
    out = b"\x89PNG\r\n\x1a\n" + png_chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, depth, color, 0, 0, int(interlace)))
    if palette is not None:
        out += png_chunk(b"PLTE", bytes(palette.reshape(-1).astype(np.uint8)))
        out += png_chunk(b"tRNS", bytes([128] * min(3, len(palette))))
    out += png_chunk(b"tEXt", b"Comment\x00decoder fixture")
    half = len(z) // 2
    out += png_chunk(b"IDAT", z[:half]) + png_chunk(b"IDAT", z[half:]) + png_chunk(b"IEND", b"")
    return out


def import_fixtures(L):
    """OBJ import, reference-order BVH and texture decoding, from the
    reference's loader, bvh_node, flat exporter and stb_image."""
    import zlib
    # OBJ assets with transforms, seeds and aspect ratios on both sides of 1.
    cases = [("gallery", "gallery/gallery.obj", dict(seed=7)),
             ("gallery_t", "gallery/gallery.obj", dict(seed=41, aspect=0.75, scale=(1.5, 0.5, 2), rotate=(15, 30, -45),
                                                        translate=(0.5, -1, 2))),
             ("cube_mirror", "cube/cube.obj", dict(seed=3, aspect=0.6, scale=(2, -1, 0.5), rotate=(30, -45, 200),
                                                   translate=(1, 2, 3))),
             ("sphere_t", "blender_sphere/sphere.obj", dict(seed=12345, aspect=16 / 9, scale=(1, 0.7, 1),
                                                            rotate=(90, 0, 0))),
             ("triangle_s0", "textured_triangle/triangle.obj", dict(seed=0, aspect=1.0)),
             ("ngon", "ngon/ngon.obj", dict(seed=2024, aspect=1.25, rotate=(0, 10, 0)))]
    for name, obj, kw in cases:
        aspect = kw.pop("aspect", 1.5)
        f = lambda v: np.asarray(v, np.float32)  # noqa: E731
        s, r, t = f(kw.get("scale", (1, 1, 1))), f(kw.get("rotate", (0, 0, 0))), f(kw.get("translate", (0, 0, 0)))
        h = L.ref_load_obj(os.path.join(ASSETS, obj).encode(), aspect, kw["seed"], s.ctypes.data, r.ctypes.data,
                           t.ctypes.data, 40.0, 0.0, 0.0, 1.0)
        if not h:
            raise RuntimeError(f"reference could not load {obj}")
        scene, cam = ref_scene(h)
        save(f"obj_{name}.npz", scene, cam, obj=np.array(obj), aspect=np.float32(aspect), seed=np.int64(kw["seed"]),
             scale=s, rotate=r, translate=t)

    # Reference-order BVH over flat inputs of 1, 2, 3 and 500 primitives.
    rng = np.random.default_rng(5)
    for name, nt, ns, seed in (("one", 1, 0, 7), ("two", 1, 1, 8), ("three", 2, 1, 9), ("soup", 500, 6, 99)):
        tri = np.zeros(nt, ptrace.TRIANGLE)
        centres = rng.normal(0, 3, (nt, 3)).astype(np.float32)
        for k in ("v0", "v1", "v2"):
            tri[k] = centres + rng.normal(0, 0.3, (nt, 3)).astype(np.float32)
        for k in ("n0", "n1", "n2"):
            n = rng.normal(0, 1, (nt, 3)).astype(np.float32)
            tri[k] = n / np.linalg.norm(n, axis=1, keepdims=True)
        for k in ("uv0", "uv1", "uv2"):
            tri[k] = np.c_[rng.random((nt, 2)), np.zeros(nt)].astype(np.float32)
        tri["smooth"] = rng.integers(0, 2, nt)
        sph = np.zeros(ns, ptrace.SPHERE)
        sph["center"] = rng.normal(0, 3, (ns, 3)).astype(np.float32)
        sph["radius"] = rng.uniform(0.2, 1, ns).astype(np.float32)
        tex = np.zeros(5, ptrace.TEXTURE)
        tex["scale"] = (1, 1, 0)
        tex["image_id"] = -1
        tex["clamp"] = 1
        tex["color"][0] = (0.2, 0.4, 0.6)
        tex["color"][1] = (4, 4, 4)
        tex[2]["type"], tex[2]["width"], tex[2]["height"], tex[2]["byte_offset"] = 1, 3, 2, 0
        tex[3]["type"], tex[3]["image_id"], tex[3]["color"] = 2, 2, (0.5, 1, 0.25)
        tex[3]["scale"], tex[3]["offset"], tex[3]["clamp"] = (2, 3, 0), (0.25, 0.5, 0), 0
        tex["color"][4] = (0.9, 0.1, 0.1)
        pixels = rng.integers(0, 256, 18).astype(np.uint8)
        mat = np.zeros(6, ptrace.MATERIAL)
        mat["ref_idx"] = 1
        mat["type"] = [0, 3, 0, 1, 2, 0]
        mat["texture_id"] = [0, 1, 3, -1, -1, 4]
        mat[3]["albedo"], mat[3]["fuzz"] = (0.8, 0.7, 0.6), 0.2
        mat[4]["ref_idx"] = 1.5
        mat[5]["texture_id"] = 4
        tri["material_id"] = rng.integers(0, 6, nt)
        sph["material_id"] = rng.integers(0, 6, ns)
        src = ptrace.FlatScene(tri, sph, np.zeros(0, ptrace.NODE), mat, tex, pixels, -1, 0)
        s, keep = oracle.scene_struct(src)
        h = L.ref_build_flat(ctypes.byref(s), seed)
        if not h:
            raise RuntimeError("reference BVH build failed")
        scene, _ = ref_scene(h)
        save(f"refbvh_{name}.npz", scene, None, in_triangles=tri, in_spheres=sph, in_materials=mat, in_textures=tex,
             in_pixels=pixels, seed=np.int64(seed))

    # PNG decoding: every colour type and bit depth, Adam7, stored / fixed /
    # dynamic DEFLATE blocks, all five row filters; expected bytes from stb_image.
    img_dir = os.path.join(HERE, "images")
    os.makedirs(img_dir, exist_ok=True)
    cases = [("gray1", 0, 1, False), ("gray2", 0, 2, True), ("gray4", 0, 4, False), ("gray8", 0, 8, True),
             ("gray16", 0, 16, False), ("graya8", 4, 8, False), ("graya16", 4, 16, True), ("rgb8", 2, 8, False),
             ("rgb8i", 2, 8, True), ("rgb16", 2, 16, False), ("rgba8", 6, 8, True), ("rgba16", 6, 16, False),
             ("pal1", 3, 1, False), ("pal2", 3, 2, True), ("pal4", 3, 4, False), ("pal8", 3, 8, True)]
    channels = {0: 1, 2: 3, 3: 1, 4: 2, 6: 4}
    expected = {}
    for k, (name, color, depth, inter) in enumerate(cases):
        w, h = int(rng.integers(1, 40)), int(rng.integers(1, 30))
        top = (1 << depth) if color != 3 else min(1 << depth, 7)
        samples = rng.integers(0, top, (h, w, channels[color]))
        palette = rng.integers(0, 256, (7, 3)) if color == 3 else None
        level, strategy = [(0, None), (6, zlib.Z_FIXED), (9, None)][k % 3]
        data = encode_png(samples, color, depth, inter, level, strategy, palette)
        path = os.path.join(img_dir, f"{name}.png")
        with open(path, "wb") as fh:
            fh.write(data)
        wi, hi = ctypes.c_int(), ctypes.c_int()
        buf = np.zeros(w * h * 3, np.uint8)
        if L.ref_load_image(path.encode(), ctypes.byref(wi), ctypes.byref(hi), buf.ctypes.data, buf.size) != 0:
            raise RuntimeError(f"stb_image rejected {name}")
        expected[name] = buf.reshape(h, w, 3)
    # A binary PPM (stb's PNM loader) and the reference's own PNG textures.
    ppm = rng.integers(0, 256, (5, 7, 3)).astype(np.uint8)
    with open(os.path.join(img_dir, "rgb.ppm"), "wb") as fh:
        fh.write(b"P6\n# fixture\n7 5\n255\n" + ppm.tobytes())
    # gray.png and rgba.png are the reference's own image fixtures (tests/fixtures/).
    for name, path in (("rgb_ppm", os.path.join(img_dir, "rgb.ppm")),
                       ("ref_gray", os.path.join(img_dir, "ref_gray.png")),
                       ("ref_rgba", os.path.join(img_dir, "ref_rgba.png")),
                       ("checker", os.path.join(ASSETS, "blender_sphere", "checker.png")),
                       ("four_colors", os.path.join(ASSETS, "textured_triangle", "materials", "four colors.png"))):
        wi, hi = ctypes.c_int(), ctypes.c_int()
        buf = np.zeros(1 << 20, np.uint8)
        if L.ref_load_image(path.encode(), ctypes.byref(wi), ctypes.byref(hi), buf.ctypes.data, buf.size) != 0:
            raise RuntimeError(f"stb_image rejected {name}")
        expected[name] = buf[:wi.value * hi.value * 3].reshape(hi.value, wi.value, 3)
    np.savez_compressed(os.path.join(HERE, "images.npz"), **expected)
    print("wrote images.npz and", len(cases) + 1, "images")


def main():
    if not oracle.ref_available():
        raise SystemExit("oracle/_ref/libptref.so is missing: run oracle/build_ref.sh first")
    L = oracle.ref()
    if sys.argv[1:] == ["import"]:
        import_fixtures(L)
        return

    # PCG streams: the reference's pixel_rng for several (seed, pixel) pairs,
    # including (42, 54) of tests/transport_tests.cpp:82.
    pairs = np.array([[42, 54], [42, 55], [7, 0], [7, 123456], [2**63 + 5, 2**40 + 3], [0, 0]], np.uint64)
    nxt = np.zeros((len(pairs), 256), np.uint32)
    uni = np.zeros((len(pairs), 256), np.float32)
    ball = np.zeros((len(pairs), 64, 3), np.float32)
    disk = np.zeros((len(pairs), 64, 3), np.float32)
    final = np.zeros(len(pairs), np.uint64)
    for i, (s, p) in enumerate(pairs):
        L.ref_rng_stream(int(s), int(p), 256, nxt[i].ctypes.data, uni[i].ctypes.data)
        st = np.zeros(1, np.uint64)
        L.ref_samplers(int(s), int(p), 64, ball[i].ctypes.data, disk[i].ctypes.data, st.ctypes.data)
        final[i] = st[0]
    rgba_in = np.array([[0, 0.25, 2], [-1, np.nan, np.inf], [0.01, 0.04, 0.09], [1, 0, 0.999999]], np.float32)
    rgba_out = np.zeros(16, np.uint8)
    L.ref_display_pixels(rgba_in.ctypes.data, len(rgba_in), rgba_out.ctypes.data)
    np.savez_compressed(os.path.join(HERE, "pcg_streams.npz"), pairs=pairs, next=nxt, uniform=uni, ball=ball,
                        disk=disk, final_state=final, rgba_in=rgba_in, rgba_out=rgba_out)
    print("wrote pcg_streams.npz")

    # Reference-exported assets (tests/cuda_tests.cpp:34-104 recipe).
    for name, obj in (("cube", "cube/cube.obj"), ("triangle", "textured_triangle/triangle.obj"),
                      ("sphere", "blender_sphere/sphere.obj")):
        scene, cam = load_asset(obj)
        rays = query_rays(scene, cam, 2048, 11)
        exp = {}
        for brute in (0, 1):
            res, boxes, prims = oracle.ref_query(scene, rays, bool(brute))
            exp[f"query_b{brute}"] = res.view(ptrace.TRACE_RESULT)
            exp[f"boxes_b{brute}"] = boxes
            exp[f"prims_b{brute}"] = prims
        lens_scene, lens_cam = load_asset(obj, aperture=0.2, t0=0.2, t1=0.8)
        exp.update(render_all(lens_scene, lens_cam, 19, 13, 7, [5, 11]))
        exp["lens_camera"] = lens_cam
        exp["rays"] = rays
        save(f"asset_{name}.npz", scene, cam, **exp)

    # Exact-tie pruning scene (tests/transport_tests.cpp:60-78).
    sph = np.zeros(3, ptrace.SPHERE)
    sph["center"] = [(0, 0, 0), (0, 0, 0), (0, 0, -8)]
    sph["radius"] = 1
    sph["material_id"] = [0, 1, 0]
    nodes = np.zeros(6, ptrace.NODE)
    spec = [((-1, -1, -9), (1, 1, 1), 0, 1, 4, -1), ((-1, -1, -1), (1, 1, 1), 0, 2, 3, -1),
            ((-1, -1, -1), (1, 1, 1), 2, -1, -1, 0), ((-1, -1, -1), (1, 1, 1), 2, -1, -1, 1),
            ((-1, -1, -9), (1, 1, -7), 0, 5, 5, -1), ((-1, -1, -9), (1, 1, -7), 2, -1, -1, 2)]
    for i, (mn, mx, ty, l, r, p) in enumerate(spec):
        nodes[i] = (mn, mx, ty, l, r, p)
    mats = np.zeros(2, ptrace.MATERIAL)
    mats["type"] = ptrace.DIFFUSE
    mats["texture_id"] = 0
    tex = np.zeros(1, ptrace.TEXTURE)
    tex["color"] = (0.5, 0.5, 0.5)
    tex["scale"] = (1, 1, 0)
    tex["clamp"] = 1
    tex["image_id"] = -1
    pruning = ptrace.FlatScene(np.zeros(0, ptrace.TRIANGLE), sph, nodes, mats, tex, np.zeros(0, np.uint8), 0, 3)
    pr = np.zeros(1, ptrace.RAY)
    pr["origin"] = (0, 0, 3)
    pr["direction"] = (0, 0, -1)
    exp = {}
    for brute in (0, 1):
        res, boxes, prims = oracle.ref_query(pruning, pr, bool(brute))
        exp[f"query_b{brute}"] = res.view(ptrace.TRACE_RESULT)
        exp[f"boxes_b{brute}"] = boxes
        exp[f"prims_b{brute}"] = prims
    save("pruning.npz", pruning, None, rays=pr, **exp)

    # Sphere scene with lens and shutter (tests/cuda_tests.cpp:105-118).
    scene, cam = ref_scene(L.ref_sphere_scene(7))
    exp = render_all(scene, cam, 31, 21, 19, [16, 48], modes=(0,), direct=(False,))
    exp.update(render_all(scene, cam, 31, 21, 19, [64], modes=(1, 2, 3, 4, 5)))
    save("spheres.npz", scene, cam, **exp)

    # Triangle light with and without an occluder (tests/lighting_tests.cpp:14-24,101-114).
    for occ in (0, 1):
        scene, cam = ref_scene(L.ref_triangle_light_scene(occ))
        exp = render_all(scene, cam, 19, 13, 11, [16], modes=(0,), direct=(True,))
        save(f"triangle_light_{occ}.npz", scene, cam, **exp)

    # Config-1 room built by this project's builder, rendered by the reference
    # (path limit 50, which the reference fixes) at a reduced 40x25.
    scene, cam = ptrace.cornell(7, 40, 25)
    exp = render_all(scene, cam, 40, 25, 7, [3, 5], modes=MODES, direct=(False, True))
    rays = query_rays(scene, cam, 1024, 5)
    res, boxes, prims = oracle.ref_query(scene, rays, False)
    save("cornell.npz", scene, cam, rays=rays, query_b0=res.view(ptrace.TRACE_RESULT), boxes_b0=boxes, prims_b0=prims,
         **exp)

    import_fixtures(L)


if __name__ == "__main__":
    main()
