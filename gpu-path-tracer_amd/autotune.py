"""Autotuner of the wavefront core: sweeps the traversal block size, the path
pool (queue capacity) and the BVH node width on a given workload and writes
the fastest pt_tuning to JSON.

The reference fixes these by hand (LaunchUtils.cuh:7-12 block sizes,
PathTracerDefaults.hpp:9 pool size, a 2-wide 40-byte BVH node); here they are
measured on the device with the workload's own batch of views.

    python gpu-path-tracer_amd/autotune.py --workload batch --out tuning.json
    python gpu-path-tracer_amd/autotune.py --workload production --out tuning.json

`apply(renderer, path)` loads a result into a renderer (scenes uploaded after
the call use its node width).
"""
import argparse
import itertools
import json
import os
import statistics
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
if HERE not in sys.path:
    sys.path.insert(0, HERE)

import ptcore  # noqa: E402
import scenes  # noqa: E402

BLOCKS = (64, 128, 256, 512)
POOLS = (1 << 23, 1 << 24, 1 << 25, 1 << 26, 1 << 27)
WIDTHS = (4, 2)
TUNED_KEYS = ("trace_block", "pool_size", "bvh_width")


def _upload(r, scene_list):
    """Uploads scene_list[0] as scene 0 and the rest with AddScene; returns ids."""
    info_ids = [0]
    r.InitializeSceneBuffers(scene_list[0])
    for s in scene_list[1:]:
        info_ids.append(r.AddScene(s))
    return info_ids


def _remove(r, ids):
    for sid in ids[1:]:
        r.RemoveScene(sid)


def autotune(r, scene_list, views, blocks=BLOCKS, pools=POOLS, widths=WIDTHS, repeats=3, log=print):
    """Times RenderBatch(views) for every (width, block, pool) point.  `views`
    are (scene index, settings, View) with indices into scene_list.  Returns
    {"best": {...}, "trials": [...]}; the renderer is left at the best point
    with the scenes uploaded at its node width."""
    base = r.get_tuning()
    trials = []
    for width in widths:
        r.set_tuning(bvh_width=width)
        t0 = time.time()
        ids = _upload(r, scene_list)
        log(f"autotune: width {width}: scenes uploaded in {time.time() - t0:.1f}s")
        items = [(ids[i], st, v) for i, st, v in views]
        for block, pool in itertools.product(blocks, pools):
            try:
                r.set_tuning(trace_block=block, pool_size=pool)
                _, st = r.RenderBatch(items)                 # warm-up: allocation and graph capture
                times = []
                for _ in range(repeats):
                    _, st = r.RenderBatch(items)
                    times.append(st["render_ms"])
                ms = statistics.median(times)
                rays = st["extension_rays"] + st["shadow_rays"]
                trials.append(dict(bvh_width=width, trace_block=block, pool_size=pool, render_ms=ms,
                                   mrays_per_s=rays / ms / 1e3, rays=rays))
                log(f"autotune: width {width} block {block:3d} pool 2^{pool.bit_length() - 1}: "
                    f"{ms:8.2f} ms  {rays / ms / 1e3:8.1f} Mrays/s")
            except ptcore.PtError as e:
                trials.append(dict(bvh_width=width, trace_block=block, pool_size=pool, error=str(e)))
                log(f"autotune: width {width} block {block} pool {pool}: {e}")
        _remove(r, ids)
    ok = [t for t in trials if "render_ms" in t]
    if not ok:
        raise RuntimeError("no tuning point ran")
    best = min(ok, key=lambda t: t["render_ms"])
    tuned = dict(base)
    tuned.update({k: best[k] for k in TUNED_KEYS})
    r.set_tuning(**tuned)
    _upload(r, scene_list)
    return {"best": {k: tuned[k] for k in tuned}, "best_trial": best, "trials": trials}


def apply(r, path):
    """Sets the tuning stored by autotune() (JSON file) on renderer r."""
    with open(path) as f:
        best = json.load(f)["best"]
    r.set_tuning(**best)
    return best


def workload(name):
    if name == "batch":
        sc, views = scenes.batch_workload()
        return sc, views, "batch: 32 scenes 10k-2M triangles, 84 views 256^2-1080p, 4-64 spp"
    if name == "production":
        s = scenes.production_scene()
        return [s], [(0, s.settings, s.view)], "production: 983,042 triangles, 1920x1080, 64 spp, depth 8"
    if name == "cornell":
        s = scenes.cornell_box()
        return [s], [(0, s.settings, s.view)], "cornell: 36 triangles, 256x256, 16 spp, depth 8"
    raise ValueError(name)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", choices=["batch", "production", "cornell"], default="production")
    ap.add_argument("--out", default="tuning.json")
    ap.add_argument("--repeats", type=int, default=3)
    ap.add_argument("--blocks", default=",".join(map(str, BLOCKS)))
    ap.add_argument("--pools", default=",".join(str(p.bit_length() - 1) for p in POOLS), help="log2 sizes")
    ap.add_argument("--widths", default=",".join(map(str, WIDTHS)))
    a = ap.parse_args()
    t0 = time.time()
    sc, views, desc = workload(a.workload)
    print(f"autotune: {desc}; workload built in {time.time() - t0:.1f}s", flush=True)
    with ptcore.Renderer() as r:
        res = autotune(r, sc, views, blocks=[int(b) for b in a.blocks.split(",")],
                       pools=[1 << int(p) for p in a.pools.split(",")], widths=[int(w) for w in a.widths.split(",")],
                       repeats=a.repeats, log=lambda m: print(m, flush=True))
    res["workload"] = desc
    with open(a.out, "w") as f:
        json.dump(res, f, indent=1)
    b = res["best_trial"]
    print(f"autotune: best width {b['bvh_width']} block {b['trace_block']} pool 2^{b['pool_size'].bit_length() - 1}: "
          f"{b['render_ms']:.2f} ms, {b['mrays_per_s']:.1f} Mrays/s -> {a.out}", flush=True)


if __name__ == "__main__":
    main()
