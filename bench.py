"""Benchmark of the plain render path on one MI355X.

Workload: 1920x1080, 64 samples per pixel, path depth 8, MIS direct lighting,
over a geodesic icosphere of 20*224^2 = 1,003,520 triangles displaced by
3-octave value noise, with a 2-triangle floor and four rectangular lights
(1,003,530 triangles in all), binned-SAH BVH.  The scene is synthetic and
seeded (seed 7).

One step = one full frame: reset(camera) + add_samples(64) through the C ABI,
inputs resident in HBM.  ``value`` is rays per second over the timed steps,
counting every closest-hit and shadow query (the reference's rays_traced).

Also reported:
  roofline     -- the path-lane kernel (every path of the frame: 4-wide
                  traversal and shading on per-lane state machines, two passes
                  per frame), HBM-bound.  achieved = algorithmic bytes of the
                  frame / HIP-event time of the frame's path-lane work on the
                  renderer stream.  Algorithmic bytes are what that traversal
                  and shading must read and write, counted by an untimed
                  counting pass over the same queries with the same 4-wide
                  traversal: 128 B per wide node visited, 48 B per leaf record
                  tested, 92 B per shaded hit (shading record 64 B + material
                  28 B), 24 B per sample (read and write of the pixel's sum),
                  16 B per pixel (PCG state in and out), and 32 B per branch /
                  36 B per leaf popped by the reference-order retraces.
                  Measured figures of the same kernel from rocprofv3 counters
                  (committed summary under profiles/): HBM bytes (traffic,
                  FETCH_SIZE + WRITE_SIZE scaled by a plain-copy calibration),
                  the measured HBM fraction, the L2 hit rate, active lanes
                  per VALU instruction and the share of SIMD cycles issuing
                  VALU instructions (valu_busy: the kernel's actual bound).
                  The reference-order figure (the reference's own
                  trace_stats: branch nodes and leaves popped by its binary
                  traversal) is reported under reference_order, and the
                  traversal's FP32 test arithmetic (12 FLOP per box test, 48
                  per triangle, 30 per sphere) against the 157.3 TFLOP/s
                  vector peak under fp32.
  cpu_baseline -- the CPU oracle (a C restatement of the reference's host
                  transport path) on every core available to the process over
                  a bounded sample of the same frame; reference_host: the
                  reference's own host path compiled from its sources
                  (oracle/_ref), labelled, at its fixed path limit of 50;
                  single_thread: the oracle on one thread, 1 spp over the
                  frame.
  denoised     -- time to relMSE <= 0.01 against a 16384-spp render of the
                  same 1080p view, for plain progressive accumulation and for a
                  low-spp render denoised by the shipped network.
"""
import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [os.path.join(HERE, "cuda-path-tracer_amd"), os.path.join(HERE, "oracle")]

WIDTH, HEIGHT, SPP, DEPTH, FREQ, SEED = 1920, 1080, 64, 8, 224, 7
DRAIN = 1 << 30  # the renderer's default: every path on the path-lane kernel
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def available_cores():
    """CPUs this process may run on: its affinity mask, limited by a cgroup
    CPU quota where one is set."""
    n = len(os.sched_getaffinity(0))
    try:
        quota, period = open("/sys/fs/cgroup/cpu.max").read().split()
        if quota != "max":
            n = min(n, max(1, int(float(quota) / float(period))))
    except (OSError, ValueError):
        pass
    return n


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(scene, camera, threads, spp):
    import oracle
    r = oracle.Renderer(scene, camera, WIDTH, HEIGHT, SEED, False, 0, True, DEPTH)
    t0 = time.perf_counter()
    r.add_samples(spp, threads=threads)
    dt = time.perf_counter() - t0
    rays = r.rays_traced()
    return {"value": rays / dt / 1e6, "unit": "Mrays/s", "cores": threads, "kind": "port", "cpu": cpu_model(),
            "sample": f"{spp} spp (the timed GPU frame has {SPP}) over the full {WIDTH}x{HEIGHT} frame of the same "
                      f"scene, camera, seed and depth ({rays} rays, {dt:.1f} s, {threads} threads)"}


def single_thread_baseline(scene, camera):
    """The oracle on one thread over one sample of the full frame (about
    5.4M rays): the per-core rate beside the all-core figure."""
    import oracle
    r = oracle.Renderer(scene, camera, WIDTH, HEIGHT, SEED, False, 0, True, DEPTH)
    t0 = time.perf_counter()
    r.add_samples(1, threads=1)
    dt = time.perf_counter() - t0
    rays = r.rays_traced()
    return {"value": rays / dt / 1e6, "unit": "Mrays/s", "cores": 1, "kind": "port",
            "sample": f"1 spp over the full {WIDTH}x{HEIGHT} frame ({rays} rays, {dt:.1f} s, 1 thread)"}


def reference_host_baseline(scene, camera, threads, rows=216):
    """The reference's own host path (portable_renderer's trace_pixel_sample,
    compiled from the reference's sources into oracle/_ref) on `rows` rows of
    the frame, 1 spp.  Its path limit is fixed at 50 (transport.h:174), so
    its rays per sample differ from the depth-8 workload: a labelled side
    figure, not the baseline."""
    import ctypes
    import numpy as np
    import oracle
    if not oracle.ref_available():
        return None
    L = oracle.ref()
    s, keep = oracle.scene_struct(scene)
    cam = np.ascontiguousarray(np.asarray(camera).reshape(()))
    y0 = HEIGHT // 2 - rows // 2
    out = np.zeros((rows * WIDTH, 3), np.float32)
    rays = ctypes.c_uint64()
    t0 = time.perf_counter()
    rc = L.ref_render_rows(ctypes.byref(s), cam.ctypes.data, WIDTH, HEIGHT, SEED, 1, 1, y0, y0 + rows, threads,
                           out.ctypes.data, ctypes.byref(rays))
    dt = time.perf_counter() - t0
    if rc != 0:
        return None
    return {"value": rays.value / dt / 1e6, "unit": "Mrays/s", "cores": threads, "kind": "reference",
            "sample": f"1 spp over rows {y0}..{y0 + rows - 1} of the same frame at the reference's fixed path limit "
                      f"of 50 ({rays.value} rays, {dt:.1f} s)"}


def profile_figures(kernel):
    """Counter figures of `kernel` from the committed summary, or {}."""
    path = os.path.join(HERE, "profiles", "path_kernel_icosphere1M_1080p_traffic.json")
    try:
        with open(path) as f:
            data = json.load(f)
    except (OSError, ValueError):
        return {}
    return data if data.get("kernel") == kernel else {}


BATCH_SIZES = [(256, 256), (512, 384), (640, 480), (800, 600), (1024, 576), (1280, 720)]
BATCH_SPP, BATCH_DEPTH = 32, 6


def batch_workload(ptrace, count, seed=SEED):
    """Seeded clustered-triangle soups, log-uniform 10k..2M triangles, each with
    its own camera and a resolution from 256x256 to 1280x720."""
    import numpy as np
    from concurrent.futures import ThreadPoolExecutor
    rng = np.random.default_rng(seed)
    tris = np.exp(rng.uniform(np.log(1e4), np.log(2e6), count)).astype(np.int64)
    sizes = [BATCH_SIZES[int(i)] for i in rng.integers(0, len(BATCH_SIZES), count)]
    def make(k):
        w, h = sizes[k]
        return ptrace.soup(int(tris[k]), seed * 1000 + k, w, h)
    workers = min(16, len(os.sched_getaffinity(0)))
    with ThreadPoolExecutor(workers) as ex:
        built = list(ex.map(make, range(count)))
    scenes = [s for s, _ in built]
    views = [(k, built[k][1], sizes[k][0], sizes[k][1], seed + k) for k in range(count)]
    return scenes, views, int(tris.sum())


def bench_build(args, ptrace):
    """The device linear BVH builder (pt_build_bvh) on the frame's scene:
    device time per build over the timed steps (HIP events around the build's
    kernels), the call's wall time (host checks, copies in and out, the depth
    walk), the host binned-SAH build of the same primitives, and per tree
    the renderer's creation (host 4-wide layout and upload) and its 16-spp
    1080p render rate."""
    scene, camera = ptrace.icosphere(FREQ, SEED, WIDTH, HEIGHT)
    arrays = (scene.triangles, scene.spheres, scene.materials, scene.textures, scene.pixels)
    n = len(scene.triangles) + len(scene.spheres)
    for _ in range(args.warmup):
        ptrace.build_bvh_device(*arrays)
    device_ms, call_ms = [], []
    for _ in range(args.steps):
        t = time.perf_counter()
        built, ms = ptrace.build_bvh_device(*arrays)
        call_ms.append((time.perf_counter() - t) * 1e3)
        device_ms.append(ms)
        log(f"device build: {ms:.3f} ms on the device, {call_ms[-1]:.1f} ms for the call")
    t = time.perf_counter()
    sah = ptrace.build_scene(*arrays, split="sah")
    host_ms = (time.perf_counter() - t) * 1e3
    rates, create_ms = {}, {}
    for name, tree in (("sah", sah), ("device", built)):
        t = time.perf_counter()
        r = ptrace.Renderer(tree, camera, WIDTH, HEIGHT, SEED, False, ptrace.DEBUG_SHADED, True, DEPTH)
        create_ms[name] = round((time.perf_counter() - t) * 1e3, 1)
        r.add_samples(1)
        r.reset(camera)
        r.add_samples(16)
        rates[name] = round(r.rays_traced() / r.kernel_milliseconds() / 1e3, 1)
        r.close()
    mean = sum(device_ms) / len(device_ms)
    print(json.dumps({
        "metric": "bvh_build_primitives_per_s", "value": round(n / (mean / 1e3), 1), "unit": "primitives/s",
        "higher_is_better": True, "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "device_ms": round(mean, 4), "call_ms": round(sum(call_ms) / len(call_ms), 2),
        "host_sah_ms": round(host_ms, 1), "host_sah_depth": sah.max_depth, "device_depth": built.max_depth,
        "render_mrays_16spp": rates, "renderer_create_ms": create_ms, "data": "synthetic",
        "config": {"workload": f"device linear BVH of the {n}-triangle 1080p icosphere frame", "primitives": n}}))
    return 0


def bench_batch(args, ptrace):
    t0 = time.perf_counter()
    scenes, views, total_tris = batch_workload(ptrace, args.scenes)
    pixels = sum(w * h for _, _, w, h, _ in views)
    log(f"batch: {len(scenes)} scenes, {total_tris} triangles, {pixels} pixels, built in {time.perf_counter() - t0:.1f}s")
    t0 = time.perf_counter()
    b = ptrace.Batch(scenes, views, direct_light=True, max_depth=BATCH_DEPTH)
    del scenes
    log(f"batch upload {time.perf_counter() - t0:.1f}s")
    tuned = None
    if args.autotune:
        t0 = time.perf_counter()
        best, rate, trials = b.autotune(probe_samples=4, passes=1)
        tuned = {"best": best, "probe_mrays": round(rate, 1), "trials": len(trials),
                 "seconds": round(time.perf_counter() - t0, 1)}
        log("autotune:", tuned)
    for _ in range(args.warmup):
        b.reset()
        b.add_samples(BATCH_SPP)
    rays = 0
    t0 = time.perf_counter()
    for _ in range(args.steps):
        b.reset()
        b.add_samples(BATCH_SPP)
        rays += b.stats()["rays"]
    elapsed = time.perf_counter() - t0
    mrays = rays / elapsed / 1e6
    st = b.last_launch_stats()
    line = {"metric": "Mrays/s, 64-scene batch", "value": round(mrays, 2), "unit": "Mrays/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(elapsed / args.steps * 1e3, 2),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
            "data": "synthetic (seeded clustered triangle soups)",
            "config": {"workload": f"batched persistent rendering, {len(views)} scenes, {total_tris} triangles, "
                                   f"{pixels} pixels, {BATCH_SPP} spp, depth {BATCH_DEPTH}, MIS",
                       "scenes": len(views), "triangles": total_tris, "pixels": pixels, "spp": BATCH_SPP,
                       "max_depth": BATCH_DEPTH, "rays_per_step": rays // args.steps,
                       "bounce_iterations": st["iterations"], "graph_replays": st["trace_launches"],
                       "drained_paths": st["drained_paths"], "tuning": b.tuning(), "autotune": tuned}}
    print(json.dumps(line), flush=True)
    return 0


def denoised_vs_plain(scene, camera, log):
    """Wall time to relMSE <= 0.01 against a 16384-spp render of the
    benchmark view: plain progressive accumulation (checkpoints from 4 to 4096
    spp) against a 1, 4 or 16 spp render denoised by the shipped network (the
    render, the denoiser and the read-back of the result to host memory).
    Each form is timed 3 times from a reset renderer; the times are medians,
    with the individual times of the two hits under `seconds_each`."""
    import denoise
    t0 = time.perf_counter()
    net = denoise.Denoiser(weights=denoise.WEIGHTS_PATH)
    res = denoise.time_to_relmse(scene, camera, WIDTH, HEIGHT, net, target=0.01, reference_spp=16384,
                                 log=lambda *a: log(*a))
    p, d = res["plain_hit"], res["denoised_hit"]
    out = {"metric": "seconds to relMSE <= 0.01 (1080p, against 16384 spp)", "reference_spp": 16384,
           "relmse_eps": 0.01,
           "plain": ({"spp": p["spp"], "seconds": round(p["seconds"], 4), "relmse": round(p["relmse"], 5),
                      "seconds_each": p["seconds_each"]} if p else None),
           "denoised": ({"spp": d["spp"], "seconds": round(d["seconds"], 4), "relmse": round(d["relmse"], 5),
                         "denoise_ms": round(d["denoise_ms"], 3), "seconds_each": d["seconds_each"]} if d else None),
           "speedup": round(res["speedup"], 2) if res["speedup"] else None,
           "denoised_by_spp": [{"spp": x["spp"], "relmse": round(x["relmse"], 5), "seconds": round(x["seconds"], 4)}
                               for x in res["denoised"]],
           "plain_by_spp": [{"spp": x["spp"], "relmse": round(x["relmse"], 5), "seconds": round(x["seconds"], 4)}
                            for x in res["plain"]],
           "wall_s": round(time.perf_counter() - t0, 1)}
    log("denoised vs plain:", out["plain"], out["denoised"], out["speedup"])
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-denoise", action="store_true", help="skip the denoised-vs-plain measurement")
    ap.add_argument("--spp", type=int, default=SPP)
    ap.add_argument("--copy-probe", action="store_true",
                    help="run only a plain 1 GiB device copy (calibrates rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    ap.add_argument("--workload", choices=["frame", "batch", "build"], default="frame",
                    help="frame: the 1080p 1M-triangle frame (default); batch: 64 scenes in one batched loop; "
                         "build: the device BVH builder on the frame's scene")
    ap.add_argument("--scenes", type=int, default=64, help="batch workload: number of scenes")
    ap.add_argument("--autotune", action="store_true", help="batch workload: run the autotuner before timing")
    args = ap.parse_args()
    if args.copy_probe:
        import torch
        n = 1 << 28
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": 4 * n, "copies": 3}))
        return 0
    if args.gpus != 1:
        print("bench.py measures one GPU: there is no multi-GPU path (pixels would be sharded by rows "
              "across GPUs with one final gather). Run with --gpus 1.", file=sys.stderr)
        return 2

    import ptrace
    ok, device = ptrace.device_available()
    if not ok:
        raise SystemExit(f"no HIP device: {device}")
    log("device:", device)
    if args.workload == "batch":
        return bench_batch(args, ptrace)
    if args.workload == "build":
        return bench_build(args, ptrace)
    t0 = time.perf_counter()
    scene, camera = ptrace.icosphere(FREQ, SEED, WIDTH, HEIGHT)
    log(f"scene: {len(scene.triangles)} triangles, {len(scene.nodes)} nodes, depth {scene.max_depth}, "
        f"built in {time.perf_counter() - t0:.1f}s")
    t0 = time.perf_counter()
    r = ptrace.Renderer(scene, camera, WIDTH, HEIGHT, SEED, False, ptrace.DEBUG_SHADED, True, DEPTH)
    log(f"upload {time.perf_counter() - t0:.2f}s")

    # Untimed counting passes over the same frame (same seed, reset), on the
    # wavefront bounce loop with no tail kernel, so that every query goes
    # through the counting trace kernel:
    # 1. the timed traversal: 4-wide nearest-first with its reference-order
    #    retraces (persistent-lane trace kernel) -- wide nodes visited, leaf
    #    records tested, closest-hit queries that hit, retrace pops;
    # 2. the reference order (binary traversal): the reference's own
    #    trace_stats (transport.h:43-85), reported as a separate figure.
    def counting_pass(exact):
        r.reset(camera)
        r.set_tuning(count_tests=True, drain_threshold=-1, exact_traversal=exact)
        r.add_samples(args.spp)
        st = r.last_launch_stats()
        ext_log, sh_log = r.iteration_log()
        return st, int(ext_log.sum()), int(sh_log.sum())
    wide_counts, wide_ext, wide_sh = counting_pass(False)
    log("counting pass (4-wide):", wide_counts)
    ref_counts, traced_ext, traced_sh = counting_pass(True)
    log("counting pass (reference order):", ref_counts)
    r.set_tuning(count_tests=False, time_kernels=True, drain_threshold=DRAIN, exact_traversal=False)

    for _ in range(args.warmup):
        r.reset(camera)
        r.add_samples(args.spp)
    rays = 0
    trace_ms = shade_ms = regen_ms = drain_ms = 0.0
    launches = splits = retraced = 0
    t0 = time.perf_counter()
    for _ in range(args.steps):
        r.reset(camera)
        r.add_samples(args.spp)
        st = r.last_launch_stats()
        rays += st["extend_rays"] + st["shadow_rays"]
        trace_ms += st["trace_ms"]
        shade_ms += st["shade_ms"]
        regen_ms += st["regen_ms"]
        drain_ms += st["drain_ms"]
        launches += st["trace_launches"]
        splits += st["split_launches"]
        retraced += st["retraced_queries"]
    elapsed = time.perf_counter() - t0
    ms_per_step = elapsed / args.steps * 1e3
    mrays = rays / elapsed / 1e6
    log(f"{args.steps} steps: {elapsed:.3f}s, {mrays:.1f} Mrays/s, path lanes {drain_ms / args.steps:.2f} ms/frame")

    # Dominant kernel: the path-lane kernel (every query of the frame).  A
    # frame runs it in two passes (first samples of every pixel, then the rest
    # in cost order): the unit is the frame's path-lane work, timed from the
    # first pass's start to the second's end (HIP events on the renderer's
    # stream; the 8-bit key sort between them included).
    pixels = WIDTH * HEIGHT
    alg = {"wide_nodes": 128 * wide_counts["box_tests"],
           "leaf_records": 48 * (wide_counts["triangle_tests"] + wide_counts["sphere_tests"]),
           "shading": 92 * wide_counts["hit_queries"],
           "accumulation": 24 * pixels * args.spp + 16 * pixels,
           "retraces": 32 * wide_counts["retrace_box_tests"] + 36 * wide_counts["retrace_primitive_tests"]}
    alg_bytes = sum(alg.values())
    passes = 2 if splits else 1
    kernel = "path_lanes_kernel<false, 1> + path_lanes_kernel<false, 2>" if splits else "path_lanes_kernel<false, 0>"
    launch_s = drain_ms / 1e3 / args.steps
    if launches:  # a tuning that left paths on the wavefront loop: the bytes cover more than the path lanes
        kernel += " (with wavefront trace launches)"
    achieved = alg_bytes / launch_s / 1e9
    prof = profile_figures(kernel)
    queries = wide_ext + wide_sh
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": prof.get("hbm_bytes_per_launch"),
                "kernel": kernel, "launches_per_unit": passes, "avg_launch_ms": round(launch_s * 1e3, 4),
                "alg_bytes_per_launch": int(alg_bytes), "alg_bytes_by_kind": alg,
                "wide_nodes_per_query": round(wide_counts["box_tests"] / queries, 2),
                "leaf_records_per_query": round((wide_counts["triangle_tests"] + wide_counts["sphere_tests"]) / queries, 2),
                "retraced_queries_per_frame": retraced // max(1, args.steps),
                "measured_hbm_frac": prof.get("measured_hbm_frac"), "l2_hit_rate": prof.get("l2_hit_rate"),
                "valu_active_lanes": prof.get("valu_active_lanes"), "valu_busy": prof.get("valu_busy"),
                "profile": prof.get("summary")}
    # Secondary figure: the traversal's FP32 arithmetic against the vector
    # FP32 peak (157.3 TFLOP/s): 12 FLOP per box test (4 per child box of a
    # wide node visited), 48 per triangle test and 30 per sphere test.
    flop = (12 * 4 * wide_counts["box_tests"] + 48 * wide_counts["triangle_tests"]
            + 30 * wide_counts["sphere_tests"])
    roofline["fp32"] = {"flop_per_launch": int(flop), "achieved": round(flop / launch_s / 1e12, 2),
                        "peak": 157.3, "unit": "TFLOP/s", "frac": round(flop / launch_s / 1e12 / 157.3, 4)}
    ref_bytes = (32 * ref_counts["box_tests"] + 36 * ref_counts["triangle_tests"] + 16 * ref_counts["sphere_tests"]
                 + 48 * traced_ext + 36 * traced_sh)
    roofline["reference_order"] = {
        "alg_bytes": int(ref_bytes), "achieved": round(ref_bytes / launch_s / 1e9, 1),
        "box_tests_per_query": round(ref_counts["box_tests"] / (traced_ext + traced_sh), 2),
        "triangle_tests_per_query": round(ref_counts["triangle_tests"] / (traced_ext + traced_sh), 2),
        "note": "the reference's binary traversal tests (32 B per branch, 36 B per triangle) over the same time; "
                "not the timed kernel's own work"}

    cpu = ref_host = None
    if not args.no_cpu:
        threads = available_cores()
        spp = max(4, min(SPP, round(SPP * threads / 16)))  # the whole frame at 16 threads (~12 s)
        log(f"CPU baseline on {threads} threads ({cpu_model()}), {spp} spp ...")
        cpu = cpu_baseline(scene, camera, threads, spp)
        log("cpu:", cpu)
        ref_host = reference_host_baseline(scene, camera, threads)
        log("reference host path:", ref_host)
        if cpu and ref_host:
            cpu["reference_host"] = ref_host
        one = single_thread_baseline(scene, camera)
        log("one thread:", one)
        cpu["single_thread"] = one

    den = None
    if not args.no_denoise:
        den = denoised_vs_plain(scene, camera, log)

    line = {"metric": "Mrays/s at 1080p, 1M tris", "value": round(mrays, 2), "unit": "Mrays/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 2),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
            "data": "synthetic (seeded displaced icosphere, seed 7)",
            "config": {"workload": "plain path tracing (path lanes), icosphere 1,003,530 tris, 1920x1080, "
                                   f"{args.spp} spp, depth {DEPTH}, MIS", "width": WIDTH, "height": HEIGHT,
                       "spp": args.spp, "max_depth": DEPTH, "triangles": int(len(scene.triangles)),
                       "rays_per_step": rays // args.steps,
                       "path_lanes_ms_per_step": round(drain_ms / args.steps, 2),
                       "wavefront_trace_launches_per_step": launches / args.steps,
                       "split_launches_per_step": splits / args.steps},
            "roofline": roofline, "cpu_baseline": cpu, "denoised": den}
    if cpu:
        line["gpu_over_cpu"] = round(mrays / cpu["value"], 1)
    print(json.dumps(line), flush=True)
    return 0



if __name__ == "__main__":
    sys.exit(main())
