"""Learned denoiser over the progressive render (include/pt_denoise.h).

Python side of the BF16 MFMA encoder-decoder in libptrace.so: the network
handle, training-set generation with this renderer (low-spp radiance plus
first-hit albedo and normal, against a high-spp target of the same view),
training, and denoising a renderer's accumulation on the device.

    python cuda-path-tracer_amd/denoise.py train --out weights.npz
    python cuda-path-tracer_amd/denoise.py evaluate --weights weights.npz

The weights trained by the project ship as ``denoiser_weights.npz`` beside
this file.
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
if HERE not in sys.path:
    sys.path.insert(0, HERE)

import ptrace  # noqa: E402

WEIGHTS_PATH = os.path.join(HERE, "denoiser_weights.npz")

PTD_FUNCTIONS = {
    "ptd_last_error": (ctypes.c_char_p, []),
    "ptd_parameter_count": (ctypes.c_int64, []),
    "ptd_create": (ctypes.c_int, [ctypes.c_uint64, ctypes.c_void_p]),
    "ptd_destroy": (None, [ctypes.c_void_p]),
    "ptd_get_parameters": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64]),
    "ptd_set_parameters": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64]),
    "ptd_forward": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                   ctypes.c_void_p]),
    "ptd_loss_and_gradients": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32,
                                              ctypes.c_int32, ctypes.c_int32, ctypes.c_void_p, ctypes.c_void_p,
                                              ctypes.c_int64]),
    "ptd_set_training_data": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32,
                                             ctypes.c_int32, ctypes.c_int32]),
    "ptd_train": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, ctypes.c_float,
                                 ctypes.c_void_p, ctypes.c_void_p]),
    "ptd_denoise_renderer": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
}

_READY = False


def lib():
    global _READY
    L = ptrace.lib()
    if not _READY:
        for name, (res, args) in PTD_FUNCTIONS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _READY = True
    return L


def _check(status):
    if status == 0:
        return
    msg = lib().ptd_last_error().decode()
    if status == ptrace.PT_EINVAL:
        raise ptrace.InvalidArgument(msg)
    raise RuntimeError(msg)


def _f32(a):
    return np.ascontiguousarray(a, np.float32)


class Denoiser:
    """The network on the device (FP32 master weights, bf16 compute)."""

    def __init__(self, seed=1, weights=None):
        h = ctypes.c_void_p()
        _check(lib().ptd_create(int(seed), ctypes.byref(h)))
        self._h = h
        if weights is not None:
            self.set_parameters(np.load(weights)["parameters"] if isinstance(weights, str) else weights)

    def __del__(self):
        if getattr(self, "_h", None):
            lib().ptd_destroy(self._h)
            self._h = None

    @staticmethod
    def parameter_count():
        return int(lib().ptd_parameter_count())

    def parameters(self):
        p = np.zeros(self.parameter_count(), np.float32)
        _check(lib().ptd_get_parameters(self._h, p.ctypes.data, p.size))
        return p

    def set_parameters(self, params):
        p = _f32(params)
        _check(lib().ptd_set_parameters(self._h, p.ctypes.data, p.size))

    def save(self, path, **meta):
        np.savez_compressed(path, parameters=self.parameters(), **meta)

    def forward(self, features):
        """Denoised radiance (n,h,w,3) of features (n,h,w,9)."""
        f = _f32(features)
        n, h, w, _ = f.shape
        out = np.zeros((n, h, w, 3), np.float32)
        _check(lib().ptd_forward(self._h, f.ctypes.data, n, h, w, out.ctypes.data))
        return out

    def loss_and_gradients(self, features, targets):
        f, t = _f32(features), _f32(targets)
        n, h, w, _ = f.shape
        loss = ctypes.c_float()
        g = np.zeros(self.parameter_count(), np.float32)
        _check(lib().ptd_loss_and_gradients(self._h, f.ctypes.data, t.ctypes.data, n, h, w, ctypes.byref(loss),
                                            g.ctypes.data, g.size))
        return float(loss.value), g

    def set_training_data(self, features, targets):
        f, t = _f32(features), _f32(targets)
        crops, h, w, _ = f.shape
        _check(lib().ptd_set_training_data(self._h, f.ctypes.data, t.ctypes.data, crops, h, w))

    def train(self, crop_ids, batch, learning_rate):
        ids = np.ascontiguousarray(crop_ids, np.int32).reshape(-1)
        steps = ids.size // batch
        losses = np.zeros(steps, np.float32)
        ms = ctypes.c_float()
        _check(lib().ptd_train(self._h, ids.ctypes.data, batch, steps, float(learning_rate), losses.ctypes.data,
                               ctypes.byref(ms)))
        return losses, float(ms.value)

    def denoise_renderer(self, renderer, out=None):
        """(W*H, 3) denoised radiance of a renderer's accumulation (into `out`
        when given, e.g. a ptrace.pinned() array) and the device milliseconds."""
        if out is None:
            out = np.zeros((renderer.pixels, 3), np.float32)
        elif out.dtype != np.float32 or out.size < 3 * renderer.pixels or not out.flags.c_contiguous:
            raise ValueError("out must be a C-contiguous float32 array of at least 3 * pixels values")
        ms = ctypes.c_float()
        _check(lib().ptd_denoise_renderer(self._h, renderer._h, out.ctypes.data, ctypes.byref(ms)))
        return out, float(ms.value)


def relmse(x, ref, eps=0.01):
    """mean over pixels and channels of (x - ref)^2 / (ref^2 + eps)."""
    x = np.asarray(x, np.float64)
    ref = np.asarray(ref, np.float64)
    return float(np.mean((x - ref) ** 2 / (ref ** 2 + eps)))


# ---------------------------------------------------------------- training data
ICOSPHERE_FREQUENCIES = (24, 40, 64, 128)


def training_view_specs(count, seed):
    """Seeded variants of the small room and the displaced icosphere scenes:
    ("cornell", scene seed) or ("icosphere", frequency, scene seed), alternating."""
    rng = np.random.default_rng(seed)
    specs = []
    for i in range(count):
        if i % 2 == 0:
            specs.append(("cornell", int(rng.integers(1, 1 << 30))))
        else:
            specs.append(("icosphere", int(rng.choice(ICOSPHERE_FREQUENCIES)), int(rng.integers(1, 1 << 30))))
    return specs


def build_view(spec, size):
    if spec[0] == "cornell":
        return ptrace.cornell(spec[1], size, size)
    return ptrace.icosphere(spec[1], spec[2], size, size)


def training_views(count, seed, size):
    """(scene, camera) pairs at size x size pixels (see training_view_specs)."""
    return [build_view(spec, size) for spec in training_view_specs(count, seed)]


def render_features(scene, camera, width, height, seed, spp, depth=8):
    """(radiance, albedo, normal) of an spp-sample render with feature buffers, each (H*W, 3)."""
    r = ptrace.Renderer(scene, camera, width, height, seed, False, ptrace.DEBUG_SHADED, True, depth)
    r.enable_aovs()
    r.add_samples(spp)
    albedo, normal = r.aovs()
    return r.image(), albedo, normal


def make_training_set(views=16, frame=512, crop=256, spp_in=(1, 4, 16), spp_target=1024, seed=11, group=64,
                      log=print, first_view=0, view_count=None):
    """Crops of rendered views: features (pairs, crop, crop, 9) and targets
    (pairs, crop, crop, 3).  Each view is cut into (frame/crop)^2 crops; each
    crop gives one pair per sample count in spp_in, all against the same
    spp_target render of the view.

    Targets: `group` views at a time in one batched renderer (many 512x512
    frames fill the GPU where one does not), spp_target samples in calls of
    at most 512.  Inputs: one renderer per view with first-hit feature
    buffers, accumulated progressively through the sample counts of spp_in
    (the 4-spp input holds the 1-spp one's sample), as a caller denoising a
    progressive render would see them.  Only views [first_view, first_view +
    view_count) of the seeded list of `views` are rendered."""
    spp_in = tuple(sorted(spp_in))
    per_side = frame // crop
    per_view = per_side * per_side
    if view_count is None:
        view_count = views - first_view
    n_pairs = view_count * per_view * len(spp_in)
    feats = np.empty((n_pairs, crop, crop, 9), np.float32)
    targets = np.empty((n_pairs, crop, crop, 3), np.float32)
    specs = training_view_specs(views, seed)
    t0 = time.perf_counter()
    target_samples = 0
    target_seconds = 0.0
    k = 0
    from concurrent.futures import ThreadPoolExecutor
    workers = max(1, min(16, len(os.sched_getaffinity(0))))
    end = first_view + view_count
    for g in range(first_view, end, group):
        ids = list(range(g, min(end, g + group)))
        with ThreadPoolExecutor(workers) as ex:
            built = list(ex.map(lambda i: build_view(specs[i], frame), ids))
        batch = ptrace.Batch([sc for sc, _ in built],
                             [(j, built[j][1], frame, frame, 5000 + ids[j]) for j in range(len(ids))],
                             direct_light=True, max_depth=8)
        tb = time.perf_counter()
        done = 0
        while done < spp_target:
            step = min(512, spp_target - done)
            batch.add_samples(step)
            done += step
        tref = [batch.image(j).reshape(frame, frame, 3) for j in range(len(ids))]
        dt = time.perf_counter() - tb
        del batch
        target_samples += len(ids) * frame * frame * spp_target
        target_seconds += dt
        for j, i in enumerate(ids):
            scene, cam = built[j]
            r = ptrace.Renderer(scene, cam, frame, frame, 1000 + 31 * i, False, ptrace.DEBUG_SHADED, True, 8)
            r.enable_aovs()
            have = 0
            for spp in spp_in:
                r.add_samples(spp - have)
                have = spp
                albedo, normal = r.aovs()
                f = np.concatenate([r.image(), albedo, normal], axis=1).reshape(frame, frame, 9)
                for y in range(0, per_side * crop, crop):
                    for x in range(0, per_side * crop, crop):
                        feats[k] = f[y:y + crop, x:x + crop]
                        targets[k] = tref[j][y:y + crop, x:x + crop]
                        k += 1
            del r
        del built, tref
        log(f"views {ids[-1] + 1 - first_view}/{view_count}: {time.perf_counter() - t0:.1f}s, targets "
            f"{target_samples / max(target_seconds, 1e-9) / 1e9:.2f} G samples/s")
    assert k == n_pairs
    return feats, targets


def train(net, feats, targets, steps=2000, batch=16, lr=3e-4, seed=3, chunk=100, log=print, schedule=None):
    """Adam steps of `batch` random crops; the learning rate decays
    exponentially from lr to lr/10 over `schedule` = (first step, total steps)
    of a schedule that may span several calls (default: this call's steps)."""
    first, total = schedule if schedule else (0, steps)
    net.set_training_data(feats, targets)
    rng = np.random.default_rng(seed + first)
    history, device_ms = [], 0.0
    for s in range(0, steps, chunk):
        n = min(chunk, steps - s)
        ids = rng.integers(0, len(feats), n * batch)
        rate = lr * (0.1 ** ((first + s) / total))
        losses, ms = net.train(ids, batch, rate)
        device_ms += ms
        history += losses.tolist()
        log(f"steps {first + s + n}/{total}: loss {np.mean(losses):.4f}, {ms / n:.2f} ms/step")
    return np.array(history), device_ms


# ---------------------------------------------------------------- evaluation
def time_to_relmse(scene, camera, width, height, net, target=0.01, reference_spp=16384,
                   plain_spp=(4, 8, 16, 24, 32, 40, 48, 56, 64, 96, 128, 256, 512, 1024, 2048, 4096),
                   denoised_spp=(1, 4, 16), seed=7, log=print, repeats=3):
    """Wall time for plain progressive accumulation and for render + denoise to
    reach relMSE <= target against a high-spp render of the same view.  Both
    forms deliver their image into the same page-locked host buffer
    (ptrace.pinned), allocated once beforehand like the device buffers.
    Each form is timed `repeats` times from a reset renderer (the same
    samples each time) and every reported time is the median, so that one
    slow wall-clock sample of a few milliseconds does not decide the ratio."""
    result = ptrace.pinned((width * height, 3))
    ref = ptrace.Renderer(scene, camera, width, height, 99991, False, ptrace.DEBUG_SHADED, True, 8)
    t0 = time.perf_counter()
    done = 0
    while done < reference_spp:
        step = min(1024, reference_spp - done)
        ref.add_samples(step)
        done += step
        log(f"reference {done}/{reference_spp} spp: {time.perf_counter() - t0:.1f}s")
    reference = ref.image()
    del ref
    plain = []
    # Each timed renderer first renders one untimed sample and is reset: its
    # first add_samples call sizes the schedule's scratch buffers (set-up, like
    # the denoiser's below), and the reset restores the seeded initial state,
    # so the timed samples are those of a fresh renderer.
    r = ptrace.Renderer(scene, camera, width, height, seed, False, ptrace.DEBUG_SHADED, True, 8)
    r.add_samples(1)
    times = {spp: [] for spp in plain_spp}
    errors = {}
    for rep in range(repeats):
        r.reset(camera)
        have, spent = 0, 0.0
        for spp in sorted(plain_spp):
            t0 = time.perf_counter()
            r.add_samples(spp - have)
            spent += time.perf_counter() - t0
            have = spp
            # The image read back to the host is part of the result, as it is
            # for the denoised form (denoise_renderer returns host memory).
            t0 = time.perf_counter()
            image = r.image(out=result)
            times[spp].append(spent + time.perf_counter() - t0)
            if rep == 0:
                errors[spp] = relmse(image, reference)
    for spp in sorted(plain_spp):
        plain.append({"spp": spp, "seconds": float(np.median(times[spp])), "relmse": errors[spp],
                      "seconds_each": [round(t, 5) for t in times[spp]]})
        log("plain", plain[-1])
    # Size the network's buffers for this frame once, untimed (set-up, like
    # the renderers' own allocations).
    with ptrace.Renderer(scene, camera, width, height, seed + 1, False, ptrace.DEBUG_SHADED, True, 8) as warm:
        warm.enable_aovs()
        warm.add_samples(1)
        net.denoise_renderer(warm, out=result)
    den = []
    for spp in sorted(denoised_spp):
        r = ptrace.Renderer(scene, camera, width, height, seed, False, ptrace.DEBUG_SHADED, True, 8)
        r.enable_aovs()
        r.add_samples(1)
        secs, device_ms = [], []
        for rep in range(repeats):
            r.reset(camera)
            t0 = time.perf_counter()
            r.add_samples(spp)
            out, ms = net.denoise_renderer(r, out=result)
            secs.append(time.perf_counter() - t0)
            device_ms.append(ms)
        den.append({"spp": spp, "seconds": float(np.median(secs)), "denoise_ms": float(np.median(device_ms)),
                    "relmse": relmse(out, reference), "noisy_relmse": relmse(r.image(), reference),
                    "seconds_each": [round(t, 5) for t in secs]})
        log("denoised", den[-1])

    def first(rows):
        ok = [row for row in rows if row["relmse"] <= target]
        return ok[0] if ok else None
    p, d = first(plain), first(den)
    return {"plain": plain, "denoised": den, "plain_hit": p, "denoised_hit": d,
            "speedup": (p["seconds"] / d["seconds"]) if p and d else None}


def main():
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("train")
    t.add_argument("--out", default=WEIGHTS_PATH)
    t.add_argument("--views", type=int, default=1024)
    t.add_argument("--frame", type=int, default=512)
    t.add_argument("--crop", type=int, default=256)
    t.add_argument("--spp-in", type=int, nargs="+", default=[1, 4, 16])
    t.add_argument("--spp-target", type=int, default=4096)
    t.add_argument("--steps", type=int, default=6000)
    t.add_argument("--group", type=int, default=64, help="views per batched target render")
    t.add_argument("--parts", type=int, default=1, help="split views and steps over this many runs")
    t.add_argument("--part", type=int, default=0)
    t.add_argument("--init", default=None, help="weights to start from (the previous part's)")
    t.add_argument("--batch", type=int, default=16)
    t.add_argument("--lr", type=float, default=3e-4)
    pr = sub.add_parser("profile", help="20 training steps of batch 16 on synthetic 256x256 crops (for rocprofv3)")
    pr.add_argument("--steps", type=int, default=20)
    inf = sub.add_parser("infer", help="render the benchmark frame at low spp and denoise it (for rocprofv3)")
    inf.add_argument("--weights", default=WEIGHTS_PATH)
    inf.add_argument("--width", type=int, default=1920)
    inf.add_argument("--height", type=int, default=1080)
    inf.add_argument("--spp", type=int, default=1)
    inf.add_argument("--repeat", type=int, default=10)
    e = sub.add_parser("evaluate")
    e.add_argument("--weights", default=WEIGHTS_PATH)
    e.add_argument("--width", type=int, default=1920)
    e.add_argument("--height", type=int, default=1080)
    e.add_argument("--reference-spp", type=int, default=16384)
    e.add_argument("--out", default=None)
    args = ap.parse_args()
    log = lambda *a: print(*a, flush=True)  # noqa: E731
    if args.cmd == "profile":
        rng = np.random.default_rng(0)
        shape = (32, 256, 256)
        feats = np.concatenate([rng.gamma(1.0, 0.5, shape + (3,)), rng.uniform(0, 1, shape + (3,)),
                                rng.normal(0, 1, shape + (3,))], -1).astype(np.float32)
        targets = rng.gamma(1.0, 0.5, shape + (3,)).astype(np.float32)
        net = Denoiser(seed=1)
        net.set_training_data(feats, targets)
        _, ms = net.train(rng.integers(0, 32, 16 * args.steps), 16, 1e-4)
        log(f"{args.steps} steps of batch 16 at 256x256: {ms / args.steps:.2f} ms/step")
        net.forward(feats[:1])
        return 0
    if args.cmd == "infer":
        net = Denoiser(weights=args.weights)
        scene, cam = ptrace.icosphere(224, 7, args.width, args.height)
        with ptrace.Renderer(scene, cam, args.width, args.height, 7, False, ptrace.DEBUG_SHADED, True, 8) as r:
            r.enable_aovs()
            t0 = time.perf_counter()
            r.add_samples(args.spp)
            log(f"render {args.spp} spp at {args.width}x{args.height}: {1e3 * (time.perf_counter() - t0):.2f} ms")
            device, wall = [], []
            for _ in range(args.repeat):
                t0 = time.perf_counter()
                _, ms = net.denoise_renderer(r)
                wall.append(1e3 * (time.perf_counter() - t0))
                device.append(ms)
            log(f"denoise: device {np.median(device):.3f} ms, wall with readback {np.median(wall):.3f} ms "
                f"(median of {args.repeat})")
        return 0
    if args.cmd == "train":
        # Part p of P renders views [p*V/P, (p+1)*V/P) of the same seeded list
        # and runs steps [p*S/P, (p+1)*S/P) of one learning-rate schedule,
        # starting from the previous part's weights (--init).
        if not 0 <= args.part < args.parts or args.views % args.parts or args.steps % args.parts:
            raise SystemExit("--part must lie in [0, --parts), and --parts must divide --views and --steps")
        v0, nv = args.part * args.views // args.parts, args.views // args.parts
        s0, ns = args.part * args.steps // args.parts, args.steps // args.parts
        feats, targets = make_training_set(args.views, args.frame, args.crop, tuple(args.spp_in), args.spp_target,
                                           group=args.group, log=log, first_view=v0, view_count=nv)
        log(f"training set part {args.part + 1}/{args.parts}: views {v0}..{v0 + nv - 1}, {len(feats)} pairs of "
            f"{args.crop}x{args.crop} ({len(feats) // len(args.spp_in)} targets at {args.spp_target} spp)")
        net = Denoiser(seed=1, weights=args.init)
        history, ms = train(net, feats, targets, ns, args.batch, args.lr, log=log, schedule=(s0, args.steps))
        del feats, targets
        if args.init:
            prior = np.load(args.init)
            if "loss" in prior.files:
                history = np.concatenate([prior["loss"], history])
            if "device_ms" in prior.files:
                ms += float(prior["device_ms"])
        net.save(args.out, loss=history, device_ms=ms, config=json.dumps(vars(args)))
        log(f"saved {args.out}: final loss {history[-50:].mean():.4f}, {ms / 1e3:.1f}s on the device")
    else:
        net = Denoiser(weights=args.weights)
        scene, cam = ptrace.icosphere(224, 7, args.width, args.height)
        res = time_to_relmse(scene, cam, args.width, args.height, net, reference_spp=args.reference_spp, log=log)
        log(json.dumps(res))
        if args.out:
            with open(args.out, "w") as f:
                json.dump(res, f, indent=1)
    return 0


if __name__ == "__main__":
    sys.exit(main())
