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
  roofline     -- path-lane kernel (every path of the frame, traversal and
                  shading on per-lane state machines; a frame of more samples
                  than the first pass runs it in two passes), HBM-bound:
                  algorithmic bytes of the frame / HIP-event time of the
                  path-lane work on the renderer's stream, against 8 TB/s.
                  Bytes per query = 32*box tests + 36*triangle tests +
                  16*sphere tests + 48 (closest hit) or + 36 (shadow), the
                  tests being the reference-order traversal's trace_stats
                  (branch nodes and leaves popped) from an untimed counting
                  pass over the same rays.  traffic =
                  rocprofv3 FETCH_SIZE + WRITE_SIZE of that work (committed
                  summary under profiles/), scaled by a plain-copy calibration.
  cpu_baseline -- the CPU oracle (a C restatement of the reference's host
                  transport path) on the host's cores over a 40-spp frame of
                  the same scene and camera (about 10 s on 16 threads).
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


def traffic_from_profiles(kernel):
    """HBM bytes per unit of `kernel` (one launch, or both passes of a split
    frame) measured by rocprofv3 --pmc (committed summary), or None."""
    path = os.path.join(HERE, "profiles", "path_kernel_icosphere1M_1080p_traffic.json")
    if not os.path.exists(path):
        return None
    try:
        with open(path) as f:
            data = json.load(f)
    except (OSError, ValueError):
        return None
    return data.get("hbm_bytes_per_launch") if data.get("kernel") == kernel else None


def cpu_baseline(scene, camera, threads, spp=1):
    import oracle
    r = oracle.Renderer(scene, camera, WIDTH, HEIGHT, SEED, False, 0, True, DEPTH)
    t0 = time.perf_counter()
    r.add_samples(spp, threads=threads)
    dt = time.perf_counter() - t0
    rays = r.rays_traced()
    return {"value": rays / dt / 1e6, "unit": "Mrays/s", "cores": threads, "kind": "port",
            "sample": f"{spp} spp over the full {WIDTH}x{HEIGHT} frame of the same scene/camera "
                      f"({rays} rays, {dt:.1f} s, {threads} threads)"}


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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--spp", type=int, default=SPP)
    ap.add_argument("--copy-probe", action="store_true",
                    help="run only a plain 1 GiB device copy (calibrates rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    ap.add_argument("--workload", choices=["frame", "batch"], default="frame",
                    help="frame: the 1080p 1M-triangle frame (default); batch: 64 scenes in one batched loop")
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
        print("bench.py measures one GPU: this build has no multi-GPU path (pixels would be sharded by rows "
              "across GPUs with one final gather; see DESIGN.md). Run with --gpus 1.", file=sys.stderr)
        return 2

    import ptrace
    ok, device = ptrace.device_available()
    if not ok:
        raise SystemExit(f"no HIP device: {device}")
    log("device:", device)
    if args.workload == "batch":
        return bench_batch(args, ptrace)
    t0 = time.perf_counter()
    scene, camera = ptrace.icosphere(FREQ, SEED, WIDTH, HEIGHT)
    log(f"scene: {len(scene.triangles)} triangles, {len(scene.nodes)} nodes, depth {scene.max_depth}, "
        f"built in {time.perf_counter() - t0:.1f}s")
    t0 = time.perf_counter()
    r = ptrace.Renderer(scene, camera, WIDTH, HEIGHT, SEED, False, ptrace.DEBUG_SHADED, True, DEPTH)
    log(f"upload {time.perf_counter() - t0:.2f}s")

    # Untimed counting pass: the same frame (same seed, reset) on the wavefront
    # bounce loop with no tail kernel and the reference-order binary traversal,
    # so every query goes through the trace kernel and its box and primitive
    # tests are the reference's own trace_stats (transport.h:43-85): branch
    # nodes and leaves popped.  These counts depend on the scene and its flat
    # hierarchy only, not on the layout or the schedule that is timed.
    r.set_tuning(count_tests=True, drain_threshold=-1, exact_traversal=True)
    r.add_samples(args.spp)
    counts = r.last_launch_stats()
    ext_log, sh_log = r.iteration_log()
    traced_ext, traced_sh = int(ext_log.sum()), int(sh_log.sum())
    r.set_tuning(count_tests=False, time_kernels=True, drain_threshold=DRAIN)
    log("counting pass:", counts, "trace-kernel queries:", traced_ext, traced_sh)

    for _ in range(args.warmup):
        r.reset(camera)
        r.add_samples(args.spp)
    rays = 0
    trace_ms = shade_ms = regen_ms = drain_ms = 0.0
    launches = splits = 0
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
    elapsed = time.perf_counter() - t0
    ms_per_step = elapsed / args.steps * 1e3
    mrays = rays / elapsed / 1e6
    log(f"{args.steps} steps: {elapsed:.3f}s, {mrays:.1f} Mrays/s, trace {trace_ms:.1f} ms, shade {shade_ms:.1f} ms")

    # Dominant kernel: the path-lane kernel (every query of the frame) by
    # default, else the wavefront trace kernel.  A split frame runs it twice
    # (first samples of every pixel, then the rest in cost order): the unit
    # is then the frame's path-lane work, timed from the first pass's start to
    # the second's end (HIP events on the renderer's stream; the 8-bit key
    # sort between them included).
    alg_bytes_frame = (32 * counts["box_tests"] + 36 * counts["triangle_tests"] + 16 * counts["sphere_tests"]
                       + 48 * traced_ext + 36 * traced_sh)
    passes = 1
    if launches == 0:
        if splits:
            kernel, passes = "path_lanes_kernel<false, 1> + path_lanes_kernel<false, 2>", 2
        else:
            kernel = "path_lanes_kernel<false, 0>"
        bytes_per_launch, launch_s = alg_bytes_frame, drain_ms / 1e3 / args.steps
    else:
        kernel, bytes_per_launch, launch_s = ("trace_lanes_kernel<false, false>", alg_bytes_frame / (launches / args.steps),
                                              trace_ms / 1e3 / launches)
    achieved = bytes_per_launch / launch_s / 1e9
    traffic = traffic_from_profiles(kernel)
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "kernel": kernel, "launches_per_unit": passes, "avg_launch_ms": round(launch_s * 1e3, 4),
                "alg_bytes_per_launch": int(bytes_per_launch),
                "box_tests_per_query": round(counts["box_tests"] / (traced_ext + traced_sh), 2),
                "triangle_tests_per_query": round(counts["triangle_tests"] / (traced_ext + traced_sh), 2)}

    cpu = None
    if not args.no_cpu:
        threads = min(16, len(os.sched_getaffinity(0)))
        log(f"CPU baseline on {threads} threads ...")
        cpu = cpu_baseline(scene, camera, threads, spp=40)
        log("cpu:", cpu)

    line = {"metric": "Mrays/s at 1080p, 1M tris", "value": round(mrays, 2), "unit": "Mrays/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 2),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
            "data": "synthetic (seeded displaced icosphere, seed 7)",
            "config": {"workload": "plain path tracing (path lanes), icosphere 1,003,530 tris, 1920x1080, "
                                   f"{args.spp} spp, depth {DEPTH}, MIS", "width": WIDTH, "height": HEIGHT,
                       "spp": args.spp, "max_depth": DEPTH, "triangles": int(len(scene.triangles)),
                       "rays_per_step": rays // args.steps, "trace_ms_per_step": round(trace_ms / args.steps, 2),
                       "shade_ms_per_step": round(shade_ms / args.steps, 2),
                       "regen_ms_per_step": round(regen_ms / args.steps, 2),
                       "drain_ms_per_step": round(drain_ms / args.steps, 2),
                       "bounce_iterations": counts["iterations"], "drained_paths": counts["drained_paths"],
                       "split_launches_per_step": splits / args.steps},
            "roofline": roofline, "cpu_baseline": cpu}
    if cpu:
        line["gpu_over_cpu"] = round(mrays / cpu["value"], 1)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
