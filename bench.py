"""Benchmark of the wavefront path-tracing core on one MI355X.

One step = one full frame of the production configuration: 12 icospheres at
subdivision 6 + ground (983,042 triangles), 1920x1080, 64 spp, depth 8, NEE,
Russian roulette from bounce 3.  The scene is built and uploaded to HBM before
the timed region.  The metric counts extension + shadow rays traced per
second.  Prints one JSON line.

    python bench.py [--steps K] [--warmup W] [--gpus 1]
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "gpu-path-tracer_amd"))

METRIC = "Mrays/s (extension+shadow) on 1 MI355X at equal relMSE; p50/p99 request latency"
HBM_PEAK_GBS = 8000.0
# algorithmic bytes of one closest-hit ray: ray record (origin|tMin, direction|tMax)
# 32 B in, hit record + queue id 17 B out, 128 B per BVH4 node, 48 B per triangle
RAY_IN_B, HIT_OUT_B, NODE_B, TRI_B = 32, 17, 128, 48


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(scene, settings, view, target_s):
    """The oracle (CPU restatement of the reference stages) on the host cores,
    on a bounded band of pixels of the same frame."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    threads = oracle.default_threads()
    t0 = time.time()
    o = oracle.OracleScene(scene)
    log(f"cpu baseline: oracle BVH built in {time.time() - t0:.1f}s, {threads} threads")
    w, h = settings["width"], settings["height"]
    centre = (h // 2) * w + w // 2
    # calibrate on growing bands until one takes 0.5 s (thread start-up no
    # longer dominates), then time one band sized to about target_s
    n = 256
    while True:
        t0 = time.perf_counter()
        o.render(settings, view, centre - n // 2, centre - n // 2 + n, threads=threads)
        dt = time.perf_counter() - t0
        if dt >= 0.5 or n >= w * h // 2:
            break
        n = min(w * h // 2, n * 4)
    n = int(max(256, min(w * h // 2, n * target_s / max(dt, 1e-3))))
    t0 = time.perf_counter()
    _, ext, sh = o.render(settings, view, centre - n // 2, centre - n // 2 + n, threads=threads)
    dt = time.perf_counter() - t0
    return {"value": round((ext + sh) / dt / 1e6, 4), "unit": "Mrays/s", "cores": threads, "kind": "port",
            "sample": f"{n} pixels at the image centre of the same frame, {settings['samples_per_pixel']} spp, "
                      f"{ext} extension + {sh} shadow rays in {dt:.1f}s"}


def calibrate_copy():
    """1 GiB device-to-device copy, three times: the counters of a run under
    rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE give the scale factor of each
    counter against a known byte count."""
    import torch
    n = 1 << 28
    a = torch.ones(n, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": n * 4, "repeats": 3}), flush=True)


def measured_traffic():
    """HBM bytes per closest-hit launch from the committed rocprofv3 --pmc
    summary of this workload (FETCH_SIZE and WRITE_SIZE scaled by the copy
    calibration), or None."""
    path = os.path.join(ROOT, "profiles", "traffic_closest_production.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", choices=["production", "cornell"], default="production")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--trace-block", type=int, default=0)
    ap.add_argument("--pool", type=int, default=0)
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a 1 GiB device-to-device copy (scale reference for FETCH_SIZE/WRITE_SIZE)")
    a = ap.parse_args()
    if a.calibrate_copy:
        calibrate_copy()
        return
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (samples would shard by "
              "pixel range with one reduction of the accumulation buffer); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import ptcore
    import scenes

    t0 = time.time()
    scene = scenes.production_scene() if a.config == "production" else scenes.cornell_box()
    settings = dict(scene.settings)
    view = scene.view
    log(f"scene {scene.name}: {scene.triangle_count} triangles, generated in {time.time() - t0:.1f}s")
    r = ptcore.Renderer()
    if a.trace_block or a.pool:
        kw = {}
        if a.trace_block:
            kw["trace_block"] = a.trace_block
        if a.pool:
            kw["pool_size"] = a.pool
        r.set_tuning(**kw)
    r.InitializeRenderingBuffers(settings["width"], settings["height"])
    t0 = time.time()
    r.InitializeSceneBuffers(scene)
    bvh = r.bvh_info()
    log(f"upload + BVH4 build {time.time() - t0:.1f}s: {bvh}")

    for i in range(a.warmup):
        st = r.RenderFrame(settings, view)
        log(f"warmup {i}: {st['render_ms']:.1f} ms, {st['extension_rays']} ext + {st['shadow_rays']} shadow rays")

    rays = 0
    dev_ms = 0.0
    t0 = time.perf_counter()
    for i in range(a.steps):
        st = r.RenderFrame(settings, view)
        rays += st["extension_rays"] + st["shadow_rays"]
        dev_ms += st["render_ms"]
    wall = time.perf_counter() - t0
    value = rays / wall / 1e6
    log(f"timed: {a.steps} frames in {wall:.3f}s wall ({dev_ms:.1f} ms device): {value:.1f} Mrays/s")

    # live per-kernel timing (hipEvents on the renderer's stream) and visit counts
    prof = r.RenderFrame(settings, view, ptcore.FRAME_PROFILE)
    inst = r.RenderFrame(settings, view, ptcore.FRAME_INSTRUMENT)
    ext = inst["extension_rays"]
    nodes_per_ray = inst["closest_nodes"] / max(1, ext)
    tris_per_ray = inst["closest_tris"] / max(1, ext)
    bytes_per_ray = RAY_IN_B + HIT_OUT_B + NODE_B * nodes_per_ray + TRI_B * tris_per_ray
    launches = max(1, prof["closest_launches"])
    bytes_per_launch = bytes_per_ray * prof["extension_rays"] / launches
    avg_launch_ms = prof["closest_ms"] / launches
    achieved = bytes_per_launch / (avg_launch_ms * 1e-3) / 1e9
    sh_nodes = inst["shadow_nodes"] / max(1, inst["shadow_rays"])
    sh_tris = inst["shadow_tris"] / max(1, inst["shadow_rays"])
    log(f"shadow rays: {sh_nodes:.2f} nodes + {sh_tris:.2f} tris per ray, "
        f"{inst['shadow_rays'] / max(1, launches):.0f} rays per launch")
    log(f"closest-hit: {launches} launches, avg {avg_launch_ms:.3f} ms, {nodes_per_ray:.2f} nodes + "
        f"{tris_per_ray:.2f} tris per ray = {bytes_per_ray:.0f} B/ray; shade {prof['shade_ms']:.1f} ms, "
        f"shadow {prof['shadow_ms']:.1f} ms, regen {prof['regen_ms']:.1f} ms, bin {prof['bin_ms']:.1f} ms, closest {prof['closest_ms']:.1f} ms")
    traffic = measured_traffic()
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": traffic["hbm_bytes_per_launch"] if traffic else None,
                "traffic_detail": traffic,
                "kernel": "k_trace<closest>", "bytes_per_ray": round(bytes_per_ray, 1),
                "nodes_per_ray": round(nodes_per_ray, 3), "tris_per_ray": round(tris_per_ray, 3),
                "avg_launch_ms": round(avg_launch_ms, 4), "rays_per_launch": round(prof["extension_rays"] / launches, 1)}

    cpu = None
    if not a.no_cpu_baseline:
        cpu = cpu_baseline(scene, settings, view, a.cpu_seconds)
        log(f"cpu baseline: {cpu}")

    out = {
        "metric": METRIC, "value": round(value, 2), "unit": "Mrays/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": round(wall / a.steps * 1e3, 3), "higher_is_better": True,
        "scaling": "strong", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"{scene.name} {settings['width']}x{settings['height']} "
                               f"{settings['samples_per_pixel']}spp depth{settings['path_depth']} NEE",
                   "triangles": scene.triangle_count, "width": settings["width"], "height": settings["height"],
                   "spp": settings["samples_per_pixel"], "path_depth": settings["path_depth"],
                   "rays_per_frame": rays // a.steps, "device_ms_per_frame": round(dev_ms / a.steps, 3)},
        "roofline": roofline, "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    r.close()


if __name__ == "__main__":
    main()
