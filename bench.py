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
BF16_MFMA_PEAK_TFLOPS = 2500.0     # dense bf16 MFMA of one MI355X
# one training record of the guide network (4 layers of 64 x 64, bf16 MFMA):
# the forward pass (4 products), the weight gradients (4) and the input deltas
# of layers 3..1 (3), 2 * 64 * 64 flops each
GUIDE_TRAIN_FLOPS_PER_RECORD = 11 * 2 * 64 * 64
# flops of the inference fused into the guided shade per diffuse/GGX entry:
# three 64 x 64 layers and the 48 output rows it evaluates (8 lobes x 5 outputs
# in three 16-row tiles)
GUIDE_INFER_FLOPS_PER_ENTRY = 2 * 64 * (3 * 64 + 48)
GUIDE_PARAMS = 16640               # fp32 partial gradient of one 64-record block
# algorithmic bytes of one closest-hit ray: ray record (origin|tMin, direction|tMax)
# 32 B in, hit record + queue id 17 B out, 128 B per BVH4 node, 48 B per triangle
RAY_IN_B, HIT_OUT_B, NODE_B, TRI_B = 32, 17, 128, 48
QNODE_B = 64        # quantized BVH4 node
Q8NODE_B = 80       # 8-wide compressed node (any hit): the 80 bytes of its 128-byte line a visit reads
# one shadow ray: its 48-byte record (origin|tMax, direction|scene, radiance|pixel)
# in; the 12-byte radiance add of an unoccluded ray is not counted
SHADOW_IN_B = 48


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


def measured_traffic(kind, node_bytes):
    """HBM bytes per launch of the closest-hit ("closest") or any-hit ("any")
    traversal from the committed rocprofv3 --pmc summary of this workload
    (FETCH_SIZE and WRITE_SIZE scaled by the copy calibration), or None when
    there is none or it was measured on another node format."""
    path = os.path.join(ROOT, "profiles", f"traffic_{kind}_production.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        t = json.load(f)
    return t if t.get("node_bytes") == node_bytes else None


# ---------------------------------------------------------------- rocprofv3 summaries
# The profiles/ summaries are written by these functions from the CSV output of
# rocprofv3 -f csv runs of this script (commands in each summary).

def _short(name):
    """Kernel name without its argument list."""
    name = name.strip()
    if name.startswith("void "):
        name = name[5:]
    depth = 0
    for i, ch in enumerate(name):
        if ch == "<":
            depth += 1
        elif ch == ">":
            depth -= 1
        elif ch == "(" and depth == 0:
            return name[:i]
    return name


def _csvs(d, suffix):
    import glob
    return sorted(glob.glob(os.path.join(d, "**", f"*{suffix}"), recursive=True))


def summarize_kernel_trace(d, command, title, bench_json=None):
    """Markdown table (calls, total, average, share) per kernel from the
    *_kernel_trace.csv of a rocprofv3 --kernel-trace run in directory d; with
    the bench line of the same run, its live launch timing, roofline and
    per-iteration queue occupancy beside it."""
    import csv
    disp = []
    for path in _csvs(d, "_kernel_trace.csv"):
        with open(path) as f:
            for rec in csv.DictReader(f):
                disp.append((int(rec["Start_Timestamp"]), int(rec["End_Timestamp"]), _short(rec["Kernel_Name"])))
    disp.sort()
    # a dispatch runs alone when no other dispatch's interval meets its own
    alone = [True] * len(disp)
    for i, (s, e, _) in enumerate(disp):
        j = i + 1
        while j < len(disp) and disp[j][0] < e:
            alone[i] = alone[j] = False
            j += 1
    rows = {}
    for (s, e, k), a in zip(disp, alone):
        dt = (e - s) * 1e-3   # us
        c = rows.setdefault(k, [0, 0.0, 0, 0.0])
        c[0] += 1
        c[1] += dt
        if a and dt >= 20.0:
            c[2] += 1
            c[3] += dt
    total = sum(v[1] for v in rows.values()) or 1.0
    out = [f"# {title}", "", "Command (MI355X, from the repo root, `export TMPDIR=/tmp`):", "", "```", command, "```", "",
           "Alone: the dispatches of at least 20 us that overlap no other dispatch (the traversals of the profiled "
           "frame, which runs its stages one after another; in the replayed bounce loop each any-hit launch "
           "overlaps the next closest-hit launch).", "",
           "| kernel | calls | total ms | avg us | % | alone calls | alone avg us |",
           "|---|---:|---:|---:|---:|---:|---:|"]
    for k, (n, t, na, ta) in sorted(rows.items(), key=lambda kv: -kv[1][1]):
        out.append(f"| `{k}` | {n} | {t / 1e3:.2f} | {t / n:.1f} | {100 * t / total:.1f} | {na} | "
                   f"{(ta / na if na else 0.0):.1f} |")
    if bench_json:
        with open(bench_json) as f:
            line = json.loads([l for l in f if l.startswith("{")][-1])
        out += ["", "The same run's bench line (hipEvents on the renderer's stream, profiled frame):", ""]
        for key in ("roofline", "roofline_closest", "roofline_any"):
            r = line.get(key)
            if r:
                out.append(f"* `{key}`: {r['kernel']}, {r['avg_launch_ms']} ms per launch of {r['rays_per_launch']:.0f} "
                           f"rays, {r['nodes_per_ray']} nodes ({r['node_bytes']} B) + {r['tris_per_ray']} triangles "
                           f"= {r['bytes_per_ray']} B per ray: {r['achieved']} GB/s algorithmic, frac {r['frac']}"
                           + (f"; measured HBM {r['hbm_achieved']} GB/s, hbm_frac {r['hbm_frac']}"
                              if r.get("hbm_achieved") else ""))
        occ = line.get("queue_occupancy")
        if occ:
            out += ["", f"Queue occupancy per wavefront iteration (pt_get_queue_occupancy; path pool "
                        f"{occ['capacity']} slots; extension and shadow rays queued against the pool; hit shares "
                        f"over the frame: {occ['share']}):", "",
                    "| iteration | extension rays / pool | shadow rays / pool |", "|---:|---:|---:|"]
            for i, (e, s) in enumerate(zip(occ["ray_fill"], occ["shadow_fill"])):
                out.append(f"| {i} | {e:.4f} | {s:.4f} |")
    return "\n".join(out) + "\n"


def _pmc_table(d, min_us=20.0):
    """{kernel: {counter: mean per dispatch, "us": mean duration}} over the
    dispatches longer than min_us of the *_counter_collection.csv in d."""
    import csv
    acc = {}
    for path in _csvs(d, "_counter_collection.csv"):
        with open(path) as f:
            for rec in csv.DictReader(f):
                us = (int(rec["End_Timestamp"]) - int(rec["Start_Timestamp"])) * 1e-3
                if us < min_us:
                    continue
                k = _short(rec["Kernel_Name"])
                key = (rec["Dispatch_Id"], rec["Counter_Name"])
                a = acc.setdefault(k, {"_disp": {}, "_us": {}})
                a["_disp"][key] = a["_disp"].get(key, 0.0) + float(rec["Counter_Value"])
                a["_us"][rec["Dispatch_Id"]] = us
    out = {}
    for k, a in acc.items():
        sums, counts = {}, {}
        for (_, c), v in a["_disp"].items():
            sums[c] = sums.get(c, 0.0) + v
            counts[c] = counts.get(c, 0) + 1
        row = {c: sums[c] / counts[c] for c in sums}
        row["us"] = sum(a["_us"].values()) / len(a["_us"])
        row["dispatches"] = len(a["_us"])
        out[k] = row
    return out


# The counter runs behind profiles/pmc_production_1080p64spp.md (one group per
# run: rocprofv3 does not split counters over passes), the default command
# block of that summary.
PMC_COMMANDS = """timeout -s KILL 240 rocprofv3 --pmc <group> -f csv -d gpurun_out/prof/pmc_<name> -o run -- python3 bench.py --steps 1 --warmup 0 --no-cpu-baseline --no-serving
#   fetch: FETCH_SIZE                 write: WRITE_SIZE
#   tcc:   TCC_HIT_sum TCC_MISS_sum TCC_EA0_RDREQ_sum TCC_EA0_RDREQ_DRAM_sum
#   sq1:   SQ_WAVES SQ_INSTS_VALU SQ_INSTS_VMEM_RD SQ_INSTS_SALU SQ_INSTS_LDS SQ_WAVE_CYCLES SQ_BUSY_CYCLES SQ_ACTIVE_INST_VALU
#   sq2:   SQ_THREAD_CYCLES_VALU SQ_ACTIVE_INST_VALU SQ_INSTS_VALU GRBM_GUI_ACTIVE
timeout -s KILL 240 rocprofv3 --pmc FETCH_SIZE -f csv -d gpurun_out/prof/pmc_cal_fetch -o run -- python3 bench.py --calibrate-copy
timeout -s KILL 240 rocprofv3 --pmc WRITE_SIZE -f csv -d gpurun_out/prof/pmc_cal_write -o run -- python3 bench.py --calibrate-copy
python3 bench.py --summarize-pmc gpurun_out/prof/pmc_cal_fetch gpurun_out/prof/pmc_cal_write gpurun_out/prof/pmc_{fetch,write,tcc,sq1,sq2} --bench-json gpurun_out/prof/bench.log --out profiles/pmc_production_1080p64spp.md"""


def summarize_pmc(cal_fetch, cal_write, dirs, bench_json, command):
    """The PMC summary of a workload: per kernel HBM bytes (FETCH_SIZE and
    WRITE_SIZE scaled by the same counters of a 1 GiB device-to-device copy),
    L2 hit rate and VALU use, from one rocprofv3 --pmc run per counter group;
    returns (markdown, {"any": traffic, "closest": traffic}) where traffic is
    the per-launch record bench.py reads into roofline.traffic."""
    command = command or PMC_COMMANDS
    copy_kib = (1 << 30) / 1024.0

    def cal(d, counter):
        t = _pmc_table(d, min_us=50.0)
        v = max(row.get(counter, 0.0) for row in t.values())
        return copy_kib / v
    fs, ws = cal(cal_fetch, "FETCH_SIZE"), cal(cal_write, "WRITE_SIZE")
    rows = {}
    for d in dirs:
        for k, row in _pmc_table(d).items():
            rows.setdefault(k, {}).update(row)
    with open(bench_json) as f:
        line = json.loads([l for l in f if l.startswith("{")][-1])
    rpl = {}
    for key in ("roofline", "roofline_closest", "roofline_any"):
        if key in line and line[key]:
            rpl["any" if "any" in line[key]["kernel"] else "closest"] = line[key]
    md = ["| kernel | avg us | HBM read MB | HBM write MB | L2 hit | VALU lane util | VALU issue / CU / cycle |",
          "|---|---:|---:|---:|---:|---:|---:|"]
    traffic = {}
    for k, r in sorted(rows.items(), key=lambda kv: -kv[1].get("us", 0) * kv[1].get("dispatches", 0)):
        rd = r.get("FETCH_SIZE", 0.0) * fs * 1024
        wr = r.get("WRITE_SIZE", 0.0) * ws * 1024
        hit, miss = r.get("TCC_HIT_sum"), r.get("TCC_MISS_sum")
        l2 = hit / (hit + miss) if hit is not None and miss is not None and hit + miss > 0 else None
        lane = (r["SQ_THREAD_CYCLES_VALU"] / (64 * r["SQ_ACTIVE_INST_VALU"])
                if r.get("SQ_THREAD_CYCLES_VALU") and r.get("SQ_ACTIVE_INST_VALU") else None)
        issue = (r["SQ_ACTIVE_INST_VALU"] / (256 * r["GRBM_GUI_ACTIVE"] / 8)
                 if r.get("SQ_ACTIVE_INST_VALU") and r.get("GRBM_GUI_ACTIVE") else None)
        fmt = lambda x, n=2: "-" if x is None else f"{x:.{n}f}"
        md.append(f"| `{k}` | {r.get('us', 0):.1f} | {rd / 1e6:.1f} | {wr / 1e6:.1f} | {fmt(l2)} | {fmt(lane)} | {fmt(issue)} |")
        kind = ("any" if "WavefrontShadowSrc" in k else "closest" if "WavefrontClosestSrc" in k else None)
        if kind and ", false, false, pt::" in k and kind in rpl:
            rays = rpl[kind]["rays_per_launch"]
            traffic[kind] = {"kernel": k, "hbm_bytes_per_launch": int(rd + wr), "hbm_read_bytes_per_launch": int(rd),
                             "hbm_write_bytes_per_launch": int(wr), "fetch_scale": round(fs, 4),
                             "write_scale": round(ws, 4), "rays_per_launch": rays,
                             "hbm_bytes_per_ray": round((rd + wr) / rays, 1),
                             "algorithmic_bytes_per_ray": rpl[kind]["bytes_per_ray"],
                             "node_bytes": rpl[kind]["node_bytes"],
                             "l2_hit": None if l2 is None else round(l2, 3),
                             "source": "profiles/pmc_production_1080p64spp.md"}
    head = ["# PMC counters: production frame (983,042 triangles, 1920x1080, 64 spp)", "",
            "One counter group per run (`bench.py --summarize-pmc` wrote this file from their CSV output):", "",
            "```", command, "```", "",
            f"Copy calibration: FETCH_SIZE x {fs:.3f}, WRITE_SIZE x {ws:.3f} KiB (a 1 GiB device-to-device copy, "
            "`bench.py --calibrate-copy`, under the same counters).", "",
            "Averages per dispatch over the dispatches longer than 20 us (counter runs serialize dispatches).", "",
            "Derived columns: HBM read = FETCH_SIZE x scale KiB, HBM write = WRITE_SIZE x scale KiB; L2 hit = "
            "TCC_HIT / (TCC_HIT + TCC_MISS); VALU lane util = SQ_THREAD_CYCLES_VALU / (64 x SQ_ACTIVE_INST_VALU); "
            "VALU issue per CU per cycle = SQ_ACTIVE_INST_VALU / (256 CUs x GRBM_GUI_ACTIVE / 8 XCDs).", ""]
    return "\n".join(head + md) + "\n", traffic


def _rays(st):
    return st["extension_rays"] + st["shadow_rays"]


def occupancy_summary(occ):
    """Queue occupancy of the profiled frame's iterations (pt_get_queue_occupancy):
    the extension-ray and shadow queues filled against the path pool, the
    material queues' share of the hits, and each iteration's fills."""
    import numpy as np
    if len(occ) == 0:
        return None
    cap = int(occ["capacity"][0])
    ext = occ["extension"].astype(np.float64)
    mat = occ["material"].sum(0).astype(np.float64)
    total = max(1.0, ext.sum())
    return {
        "capacity": cap, "iterations": int(len(occ)),
        "ray_fill_mean": round(float(ext.mean() / cap), 4), "ray_fill_min": round(float(ext.min() / cap), 4),
        "shadow_fill_mean": round(float(occ["shadow"].mean() / cap), 4),
        "share": {"diffuse": round(float(mat[0] / total), 4), "ggx": round(float(mat[1] / total), 4),
                  "mirror": round(float(mat[2] / total), 4), "emissive": round(float(mat[3] / total), 4),
                  "missed": round(float(occ["missed"].sum()) / total, 4)},
        "ray_fill": [round(float(x) / cap, 4) for x in ext],
        "shadow_fill": [round(float(x) / cap, 4) for x in occ["shadow"]],
    }


def sequential_then_batched(r, items, steps, log_prefix):
    """Device time of rendering `items` one view at a time (plain rendering,
    the reference's one-request model) and all in one batch.  Returns
    (sequential Mrays/s, batched Mrays/s, batched ms, rays, max relative
    difference of a few views batched vs alone)."""
    import numpy as np
    alone, seq_ms, seq_rays = [], 0.0, 0
    per_view = []
    for it in items:                              # warm-up: allocation, graphs
        r.RenderBatch([it])
    for k, it in enumerate(items):
        (hdr,), st = r.RenderBatch([it])
        seq_ms += st["render_ms"]
        seq_rays += _rays(st)
        per_view.append((st["render_ms"], _rays(st)))
        if k < 6:
            alone.append(hdr)
    r.RenderBatch(items)
    bms, brays = [], 0
    for _ in range(steps):
        hdrs, st = r.RenderBatch(items)
        bms.append(st["render_ms"])
        brays = _rays(st)
    bms.sort()
    b_ms = bms[len(bms) // 2]
    diff = max(float(np.max(np.abs(h - a) / (np.abs(a) + 1e-3))) for h, a in zip(hdrs, alone))
    log(f"{log_prefix}: sequential {seq_ms:.1f} ms ({seq_rays} rays), batched {b_ms:.1f} ms ({brays} rays), "
        f"max rel diff batched vs alone {diff:.2e}")
    return seq_rays / seq_ms / 1e3, brays / b_ms / 1e3, b_ms, brays, diff, per_view


# Size classes of a view by its sample count (pixels x spp after rounding to a
# square): below 2^21 a view cannot fill the 2^25-slot path pool even for one
# bounce; above 2^25 it fills the pool for several bounces by itself.
BATCH_CLASSES = (("small", 0, 1 << 21), ("medium", 1 << 21, 1 << 25), ("large", 1 << 25, 1 << 62))


def view_samples(st):
    import math
    return st["width"] * st["height"] * int(math.isqrt(st["samples_per_pixel"])) ** 2


def batch_classes(r, items, per_view, steps):
    """Per size class: the views rendered one at a time (device time, from the
    sequential pass) against the class's views batched alone, with the fixed
    cost of an alone render (the time of its smallest views)."""
    out = []
    for name, lo, hi in BATCH_CLASSES:
        ks = [k for k, (_, st, _) in enumerate(items) if lo <= view_samples(st) < hi]
        if not ks:
            continue
        cls = [items[k] for k in ks]
        seq_ms = sum(per_view[k][0] for k in ks)
        seq_rays = sum(per_view[k][1] for k in ks)
        r.RenderBatch(cls)
        ms = []
        for _ in range(max(1, steps)):
            _, st = r.RenderBatch(cls)
            ms.append(st["render_ms"])
            brays = _rays(st)
        ms.sort()
        b_ms = ms[len(ms) // 2]
        row = {"class": name, "views": len(ks), "samples": int(sum(view_samples(items[k][1]) for k in ks)),
               "alone_ms": round(seq_ms, 2), "batched_ms": round(b_ms, 2),
               "alone_mrays_s": round(seq_rays / seq_ms / 1e3, 1), "batched_mrays_s": round(brays / b_ms / 1e3, 1),
               "speedup": round(seq_ms / b_ms * brays / seq_rays, 3),
               "alone_ms_per_view_min": round(min(per_view[k][0] for k in ks), 3)}
        log(f"batch class {name}: {row['views']} views, {row['samples']} samples: alone {seq_ms:.1f} ms "
            f"({row['alone_mrays_s']} Mrays/s), batched {b_ms:.1f} ms ({row['batched_mrays_s']} Mrays/s), "
            f"{row['speedup']}x")
        out.append(row)
    return out


def serve_run(r, requests, rate_hz, seed):
    """Submits `requests` at Poisson arrival times of mean rate `rate_hz`
    (all at once when rate_hz is None) to a fresh server; returns latencies
    (ms), server stats and the wall time of the run."""
    import numpy as np
    rng = np.random.default_rng(seed)
    r.ServerStart()
    t0 = time.perf_counter()
    at = 0.0
    ids = []
    for sid, st, v in requests:
        if rate_hz:
            at += rng.exponential(1.0 / rate_hz)
            delay = at - (time.perf_counter() - t0)
            if delay > 0:
                time.sleep(delay)
        ids.append(r.ServerSubmit(sid, st, v, hdr=False))
    infos = [r.ServerWait(i, 600000) for i in ids]
    stats = r.ServerStop()
    bad = [i for i in infos if i["state"] != 2]
    if bad:
        raise RuntimeError(f"{len(bad)} requests did not finish: {bad[:2]}")
    lat = np.array([i["done_ms"] - i["submit_ms"] for i in infos])
    return lat, stats


def serving_summary(r, requests, rates, seed):
    """Latency percentiles and throughput at each arrival rate, plus the
    saturated run (all requests at once) and plain sequential rendering of
    the same requests.  The server's rate is over its wall-clock busy time, so
    the sequential rate is over wall-clock time as well: each request rendered
    alone and its RGBA8 image copied to the host, as the server delivers it
    (the device time of the sequential render loops is reported beside it)."""
    import numpy as np
    seq_ms, seq_dev_ms, seq_rays = 0.0, 0.0, 0
    for it in requests[:2]:
        r.RenderBatch([it], rgba=True, hdr=False)
    for it in requests:
        t0 = time.perf_counter()
        _, st, _ = r.RenderBatch([it], rgba=True, hdr=False)
        seq_ms += (time.perf_counter() - t0) * 1e3
        seq_dev_ms += st["render_ms"]
        seq_rays += _rays(st)
    seq = seq_rays / seq_ms / 1e3
    seq_dev = seq_rays / seq_dev_ms / 1e3
    serve_run(r, requests[:4], None, seed)      # warm-up: pool, arena, graph
    lat, stats = serve_run(r, requests, None, seed)
    sat = _rays(stats) / stats["busy_ms"] / 1e3
    log(f"serving saturated: {len(requests)} requests, {stats['busy_ms']:.0f} ms busy, {sat:.1f} Mrays/s "
        f"(sequential plain {seq:.1f} Mrays/s, {seq_ms:.0f} ms wall; {seq_dev:.1f} Mrays/s over its "
        f"{seq_dev_ms:.0f} ms of device time); latency p50 {np.percentile(lat, 50):.0f} "
        f"p99 {np.percentile(lat, 99):.0f} ms")
    out = {"requests": len(requests), "sequential_mrays_s": round(seq, 1),
           "sequential_device_mrays_s": round(seq_dev, 1), "saturated_mrays_s": round(sat, 1),
           "speedup_vs_sequential": round(sat / seq, 3), "saturated_p50_ms": round(float(np.percentile(lat, 50)), 1),
           "saturated_p99_ms": round(float(np.percentile(lat, 99)), 1), "rates": []}
    sat_rate = len(requests) / (stats["busy_ms"] / 1e3)
    for frac in rates:
        rate = frac * sat_rate
        lat, stats = serve_run(r, requests, rate, seed + 1)
        mr = _rays(stats) / stats["wall_ms"] / 1e3
        p50, p99 = float(np.percentile(lat, 50)), float(np.percentile(lat, 99))
        log(f"serving at {rate:.2f} req/s ({frac:.0%} of saturation): p50 {p50:.0f} ms, p99 {p99:.0f} ms, "
            f"{mr:.1f} Mrays/s over the run")
        out["rates"].append({"rate_hz": round(rate, 3), "load": frac, "p50_ms": round(p50, 1), "p99_ms": round(p99, 1),
                             "mrays_s": round(mr, 1)})
    return out


def serving_workload(scales, count, seed, resolutions=None, spp=None):
    """Resident config-1 rooms of the given triangle counts and `count`
    seeded requests over them."""
    import numpy as np
    import scenes
    rooms = [scenes.cornell_scaled(t) for t in scales]
    rng = np.random.default_rng(seed)
    kw = {}
    if resolutions:
        kw["resolutions"] = resolutions
    if spp:
        kw["spp"] = spp
    reqs = [scenes.serve_request(rng, len(rooms), **kw) for _ in range(count)]
    return rooms, reqs


def upload_all(r, scene_list):
    ids = [0]
    r.InitializeSceneBuffers(scene_list[0])
    for s in scene_list[1:]:
        ids.append(r.AddScene(s))
    return ids


def run_batch(a):
    """Config 3: 32 scenes, 1-4 views each, in one wavefront vs one view at a time."""
    import ptcore
    import scenes
    t0 = time.time()
    sc, views = scenes.batch_workload()
    log(f"batch workload: {len(sc)} scenes, {sum(s.triangle_count for s in sc)} triangles, {len(views)} views, "
        f"built in {time.time() - t0:.1f}s")
    r = ptcore.Renderer()
    if a.tuning:
        import autotune
        log(f"tuning {autotune.apply(r, a.tuning)}")
    t0 = time.time()
    ids = upload_all(r, sc)
    log(f"scenes uploaded in {time.time() - t0:.1f}s")
    items = [(ids[i], st, v) for i, st, v in views]
    seq, bat, b_ms, brays, diff, per_view = sequential_then_batched(r, items, a.steps, "batch")
    classes = batch_classes(r, items, per_view, a.steps)
    out = {"metric": "Mrays/s (extension+shadow), 84 views of 32 scenes in one batched wavefront",
           "value": round(bat, 2), "unit": "Mrays/s", "n_gpus": 1, "steps": a.steps, "warmup": 1,
           "ms_per_step": round(b_ms, 3), "higher_is_better": True, "scaling": "strong", "vs_baseline": None,
           "dtype": "f32", "data": "synthetic",
           "config": {"workload": "batch: 32 scenes 10k-2M triangles (soups, sphere clusters), 84 views "
                                  "256x256-1920x1080, 4-64 spp, depth 8, NEE",
                      "triangles": sum(s.triangle_count for s in sc), "views": len(views), "rays": brays,
                      "tuning": r.get_tuning()},
           "sequential_mrays_s": round(seq, 2), "speedup_vs_sequential": round(bat / seq, 3),
           "max_rel_diff_batched_vs_alone": diff, "classes": classes}
    print(json.dumps(out), flush=True)
    r.close()


def run_serve(a):
    """Config 4: 8 resident rooms of 1-20M triangles (scaled by --serve-scale),
    Poisson requests of 512^2-3840x2160 and 4-256 spp at swept rates."""
    import numpy as np
    import ptcore
    sizes = np.exp(np.linspace(np.log(1e6), np.log(20e6), 8)) * a.serve_scale
    t0 = time.time()
    rooms, reqs = serving_workload([int(x) for x in sizes], a.serve_requests, 4)
    log(f"serving workload: {[s.triangle_count for s in rooms]} triangles, {len(reqs)} requests, "
        f"built in {time.time() - t0:.1f}s")
    r = ptcore.Renderer()
    t0 = time.time()
    upload_all(r, rooms)
    log(f"resident set uploaded in {time.time() - t0:.1f}s")
    del rooms
    res = serving_summary(r, reqs, (0.5, 0.8), 11)
    top = res["rates"][-1]
    out = {"metric": "Mrays/s (extension+shadow) of the serving loop; p50/p99 request latency",
           "value": res["saturated_mrays_s"], "unit": "Mrays/s", "n_gpus": 1, "steps": len(reqs), "warmup": 4,
           "ms_per_step": None, "higher_is_better": True, "scaling": "strong", "vs_baseline": None,
           "dtype": "f32", "data": "synthetic",
           "config": {"workload": f"serve: 8 resident rooms of {int(sizes[0])}-{int(sizes[-1])} triangles, "
                                  f"{len(reqs)} Poisson requests 512x512-3840x2160, 4-256 spp, depth 8"},
           "p50_ms": top["p50_ms"], "p99_ms": top["p99_ms"], "serving": res}
    print(json.dumps(out), flush=True)
    r.close()


def run_ingest(a):
    """Scene ingestion: the host's binned-SAH build against the GPU's linear
    BVH, on the production scene (983,042 triangles) and a config-1 room of
    20M triangles: scene upload + BVH build wall time, tree size, and the
    production frame's rays/s over each tree."""
    import ptcore
    import scenes
    t0 = time.time()
    prod = scenes.production_scene()
    room = scenes.cornell_scaled(int(a.ingest_triangles))
    log(f"ingest: scenes of {prod.triangle_count} and {room.triangle_count} triangles generated in "
        f"{time.time() - t0:.1f}s")
    r = ptcore.Renderer()
    st = dict(prod.settings)
    r.InitializeRenderingBuffers(st["width"], st["height"])
    runs = []
    for name, scene in (("spheres_983042", prod), (f"room_{room.triangle_count}", room)):
        for builder, label in ((0, "host binned SAH"), (1, "GPU linear BVH")):
            r.set_tuning(bvh_builder=builder)
            r.InitializeSceneBuffers(prod)                      # warm: allocations, first-use costs
            t0 = time.perf_counter()
            r.InitializeSceneBuffers(scene)
            ms = (time.perf_counter() - t0) * 1e3
            info = r.bvh_info()
            entry = {"scene": name, "triangles": scene.triangle_count, "builder": label,
                     "upload_build_ms": round(ms, 1), "nodes": info["node_count"], "depth": info["max_depth"]}
            if scene is prod:
                r.RenderFrame(st, prod.view)
                rays, dev = 0, 0.0
                for _ in range(2):
                    f = r.RenderFrame(st, prod.view)
                    rays += _rays(f)
                    dev += f["render_ms"]
                entry["frame_mrays_s"] = round(rays / dev / 1e3, 1)
            log(f"ingest: {name} with the {label}: upload + build {ms:.0f} ms, {info['node_count']} nodes, depth "
                f"{info['max_depth']}" + (f", production frame {entry['frame_mrays_s']} Mrays/s"
                                          if "frame_mrays_s" in entry else ""))
            runs.append(entry)
    big = [e for e in runs if e["scene"] != "spheres_983042"]
    out = {"metric": f"scene upload + BVH build time, {room.triangle_count}-triangle room (GPU builder)",
           "value": big[1]["upload_build_ms"], "unit": "ms", "n_gpus": 1, "steps": 1, "warmup": 1,
           "ms_per_step": big[1]["upload_build_ms"], "higher_is_better": False, "scaling": "strong",
           "vs_baseline": None, "dtype": "f32", "data": "synthetic",
           "config": {"workload": f"ingest: production scene and a {room.triangle_count}-triangle config-1 room"},
           "host_builder_ms": big[0]["upload_build_ms"], "runs": runs}
    print(json.dumps(out), flush=True)
    r.close()


def run_guide(a):
    """Config 5: the 200k-triangle two-room scene lit through a 0.5 m gap,
    1280x720, online-trained guiding (one-sample MIS with the BSDF) against
    plain BSDF sampling at equal samples; the time to equal error is the
    ratio of error x time (Monte Carlo error falls as 1/time)."""
    import numpy as np
    import ptcore
    import scenes
    room = scenes.guiding_room(200_000)
    w, h = 1280, 720
    r = ptcore.Renderer()
    r.InitializeRenderingBuffers(w, h)
    r.InitializeSceneBuffers(room)
    out_cfgs = []
    for nee in (1, 0):
        base = dict(width=w, height=h, path_depth=8, use_nee=nee, rr_start_depth=3)
        t0 = time.time()
        ref_st, _ = r.RenderGuided(dict(base, samples_per_pixel=a.guide_ref_spp, seed=99), room.view, passes=4,
                                   guide_fraction=0.0)
        ref = r.GetFrameHdr(w, h)[..., :3].copy()
        log(f"guide: NEE {nee} reference {a.guide_ref_spp} spp in {time.time() - t0:.1f}s")
        m2 = float(ref.mean()) ** 2
        spp = a.guide_spp
        st = dict(base, samples_per_pixel=spp, seed=7)
        r.RenderGuided(st, room.view, passes=4, guide_fraction=0.0)            # warm-up
        sp, _ = r.RenderGuided(st, room.view, passes=4, guide_fraction=0.0)
        e_p = float(np.mean((r.GetFrameHdr(w, h)[..., :3] - ref) ** 2)) / m2
        sg, gs = r.RenderGuided(st, room.view, passes=4, train_steps=a.guide_steps, batch=16384, learning_rate=3e-3,
                                guide_fraction=0.5, reset=True, seed=1)
        e_g = float(np.mean((r.GetFrameHdr(w, h)[..., :3] - ref) ** 2)) / m2
        ratio = e_g * sg["render_ms"] / (e_p * sp["render_ms"])
        log(f"guide: NEE {nee} {spp} spp: plain {sp['render_ms']:.1f} ms nMSE {e_p:.4g}; guided {sg['render_ms']:.1f} ms "
            f"(train {gs['train_ms']:.1f} ms, {gs['steps']} Adam steps, {gs['records']} records) nMSE {e_g:.4g}; "
            f"time to equal error {ratio:.3f}x")
        # the next frame of a progressive render: other samples, the network kept
        # from the frame above and refined with a quarter of the Adam steps (no
        # plain first pass)
        st2 = dict(st, seed=8)
        sp2, _ = r.RenderGuided(st2, room.view, passes=4, guide_fraction=0.0)
        e_p2 = float(np.mean((r.GetFrameHdr(w, h)[..., :3] - ref) ** 2)) / m2
        sw, gw = r.RenderGuided(st2, room.view, passes=4, train_steps=max(1, a.guide_steps // 4), batch=16384,
                                learning_rate=3e-3, guide_fraction=0.5, reset=False, seed=2)
        e_w = float(np.mean((r.GetFrameHdr(w, h)[..., :3] - ref) ** 2)) / m2
        ratio_w = e_w * sw["render_ms"] / (e_p2 * sp2["render_ms"])
        log(f"guide: NEE {nee} next frame, network kept: guided {sw['render_ms']:.1f} ms (train {gw['train_ms']:.1f} ms) "
            f"nMSE {e_w:.4g} vs plain {sp2['render_ms']:.1f} ms nMSE {e_p2:.4g}; time to equal error {ratio_w:.3f}x")
        # the training kernel against the bf16 MFMA peak: one launch is a
        # forward + backward over `batch` records; each 64-record block also
        # writes its fp32 partial gradient, which the Adam step sums
        batch = 16384
        us = gs["train_kernel_us"]
        tflops = GUIDE_TRAIN_FLOPS_PER_RECORD * batch / (us * 1e-6) / 1e12 if us else 0.0
        hbm = (batch // 64 * GUIDE_PARAMS * 4 + batch * 96) / (us * 1e-6) / 1e9 if us else 0.0
        train_roof = {"bound": "mfma", "kernel": "k_guide_train", "achieved": round(tflops, 2),
                      "peak": BF16_MFMA_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": round(tflops / BF16_MFMA_PEAK_TFLOPS, 4),
                      "flops_per_record": GUIDE_TRAIN_FLOPS_PER_RECORD, "records_per_launch": batch,
                      "avg_launch_us": round(us, 2), "hbm_gbs": round(hbm, 1)}
        log(f"guide: k_guide_train {us:.1f} us per {batch}-record launch = {tflops:.1f} TFLOP/s bf16 "
            f"({100 * tflops / BF16_MFMA_PEAK_TFLOPS:.2f} % of peak), {hbm:.0f} GB/s of partial gradients and records")
        # the inference fused into the guided shade: its MLP flops over all the
        # time guiding adds to the frame (features, inference, guided sampling
        # and density, records), so a lower bound of the inference's MFMA rate
        t_guide = sg["render_ms"] - gs["train_ms"] - sp["render_ms"]
        infer_flops = GUIDE_INFER_FLOPS_PER_ENTRY * gs["shaded_entries"]
        infer_tflops = infer_flops / (t_guide * 1e-3) / 1e12 if t_guide > 0 else 0.0
        infer_roof = {"bound": "mfma", "kernel": "k_shade<false, true> (fused MLP inference)",
                      "achieved": round(infer_tflops, 1), "peak": BF16_MFMA_PEAK_TFLOPS, "unit": "TFLOP/s",
                      "frac": round(infer_tflops / BF16_MFMA_PEAK_TFLOPS, 4),
                      "flops_per_entry": GUIDE_INFER_FLOPS_PER_ENTRY, "entries": gs["shaded_entries"],
                      "time_ms": round(t_guide, 2),
                      "time_basis": "guided frame - training - plain frame: all the time guiding adds"}
        log(f"guide: fused inference {gs['shaded_entries']} entries x {GUIDE_INFER_FLOPS_PER_ENTRY} flops in the "
            f"{t_guide:.1f} ms guiding adds = {infer_tflops:.0f} TFLOP/s bf16 "
            f"({100 * infer_tflops / BF16_MFMA_PEAK_TFLOPS:.1f} % of peak, a lower bound)")
        out_cfgs.append({"use_nee": nee, "spp": spp, "plain_ms": round(sp["render_ms"], 2), "plain_nmse": e_p,
                         "roofline_train": train_roof, "roofline_infer": infer_roof,
                         "guided_ms": round(sg["render_ms"], 2), "guided_train_ms": round(gs["train_ms"], 2),
                         "guided_nmse": e_g, "error_ratio_equal_samples": round(e_g / e_p, 4),
                         "time_to_equal_error_ratio": round(ratio, 4),
                         "next_frame_network_kept": {"guided_ms": round(sw["render_ms"], 2),
                                                     "guided_train_ms": round(gw["train_ms"], 2), "guided_nmse": e_w,
                                                     "plain_ms": round(sp2["render_ms"], 2), "plain_nmse": e_p2,
                                                     "time_to_equal_error_ratio": round(ratio_w, 4)},
                         "plain_mrays_s": round(_rays(sp) / sp["render_ms"] / 1e3, 1),
                         "guided_mrays_s": round(_rays(sg) / sg["render_ms"] / 1e3, 1)})
    out = {"metric": "time to equal error, guided / plain BSDF sampling (lower is better)",
           "value": out_cfgs[0]["time_to_equal_error_ratio"], "unit": "x", "n_gpus": 1, "steps": 1, "warmup": 1,
           "ms_per_step": out_cfgs[0]["guided_ms"], "higher_is_better": False, "scaling": "strong",
           "vs_baseline": None, "dtype": "bf16 MLP / f32 render", "data": "synthetic",
           "config": {"workload": f"guide: two rooms, {room.triangle_count} triangles, light behind a wall with a "
                                  f"0.5 m gap, 1280x720, {a.guide_spp} spp in 4 passes, depth 8; nMSE against "
                                  f"{a.guide_ref_spp} spp",
                      "network": "MLP 64-64-64-64 BF16 (FP32 accumulate), 8 vMF lobes, one-sample MIS alpha 0.5"},
           "roofline": out_cfgs[0]["roofline_train"], "runs": out_cfgs}
    print(json.dumps(out), flush=True)
    r.close()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", choices=["production", "cornell", "batch", "serve", "guide", "ingest"],
                    default="production")
    ap.add_argument("--ingest-triangles", type=float, default=20e6, help="room size of --config ingest")
    ap.add_argument("--guide-spp", type=int, default=256)
    ap.add_argument("--guide-ref-spp", type=int, default=4096)
    ap.add_argument("--guide-steps", type=int, default=256)
    ap.add_argument("--tuning", default=None, help="JSON written by autotune.py")
    ap.add_argument("--serve-scale", type=float, default=1.0, help="scale of the serving rooms' triangle counts")
    ap.add_argument("--serve-requests", type=int, default=48)
    ap.add_argument("--no-serving", action="store_true", help="skip the serving section of the production line")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--trace-block", type=int, default=0)
    ap.add_argument("--pool", type=int, default=0)
    ap.add_argument("--quantized", type=int, choices=range(16), default=None,
                    help="pt_tuning.bvh_quantized: bit 0 closest hit, bit 1 any hit on 64-byte quantized nodes, "
                         "bit 2 any hit on the compressed 8-wide nodes, bit 3 (with bit 2) closest hit on them "
                         "too (default: the library's)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a 1 GiB device-to-device copy (scale reference for FETCH_SIZE/WRITE_SIZE)")
    ap.add_argument("--summarize-kernel-trace", metavar="DIR", default=None,
                    help="write a markdown summary of the rocprofv3 --kernel-trace CSV output in DIR to --out")
    ap.add_argument("--summarize-pmc", nargs="+", metavar="DIR", default=None,
                    help="CAL_FETCH CAL_WRITE RUN_DIR... [--bench-json FILE]: PMC summary to --out and "
                         "profiles/traffic_{any,closest}_production.json")
    ap.add_argument("--bench-json", default=None, help="the bench.py output line of the profiled runs")
    ap.add_argument("--command", default="", help="the profiling command, recorded in the summary")
    ap.add_argument("--title", default="Kernel trace")
    ap.add_argument("--out", default=None)
    a = ap.parse_args()
    if a.calibrate_copy:
        calibrate_copy()
        return
    if a.summarize_kernel_trace:
        text = summarize_kernel_trace(a.summarize_kernel_trace, a.command, a.title, a.bench_json)
        if a.out:
            with open(a.out, "w") as f:
                f.write(text)
        print(text)
        return
    if a.summarize_pmc:
        md, traffic = summarize_pmc(a.summarize_pmc[0], a.summarize_pmc[1], a.summarize_pmc[2:], a.bench_json,
                                    a.command)
        if a.out:
            with open(a.out, "w") as f:
                f.write(md)
        for kind, t in traffic.items():
            with open(os.path.join(ROOT, "profiles", f"traffic_{kind}_production.json"), "w") as f:
                json.dump(t, f, indent=1)
        print(md)
        print(json.dumps(traffic, indent=1))
        return
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (samples would shard by "
              "pixel range with one reduction of the accumulation buffer); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    if a.config == "batch":
        run_batch(a)
        return
    if a.config == "serve":
        run_serve(a)
        return
    if a.config == "guide":
        run_guide(a)
        return
    if a.config == "ingest":
        run_ingest(a)
        return
    import ptcore
    import scenes

    t0 = time.time()
    scene = scenes.production_scene() if a.config == "production" else scenes.cornell_box()
    settings = dict(scene.settings)
    view = scene.view
    log(f"scene {scene.name}: {scene.triangle_count} triangles, generated in {time.time() - t0:.1f}s")
    r = ptcore.Renderer()
    if a.trace_block or a.pool or a.quantized is not None:
        kw = {}
        if a.trace_block:
            kw["trace_block"] = a.trace_block
        if a.pool:
            kw["pool_size"] = a.pool
        if a.quantized is not None:
            kw["bvh_quantized"] = a.quantized
        r.set_tuning(**kw)
    quant = r.get_tuning()["bvh_quantized"]
    r.InitializeRenderingBuffers(settings["width"], settings["height"])
    t0 = time.time()
    r.InitializeSceneBuffers(scene)
    bvh = r.bvh_info()
    # closest-hit node format: the 8-wide tree with bit 3, else quantized or fp32 BVH4
    node_b = Q8NODE_B if (quant & 8 and bvh["bvh8_node_count"]) else QNODE_B if quant & 1 else NODE_B
    # any-hit node format: the 8-wide tree where it was built (4-wide host build)
    node_b_any = Q8NODE_B if bvh["bvh8_node_count"] else QNODE_B if quant & 2 else NODE_B
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
    occupancy = occupancy_summary(r.queue_occupancy())
    if occupancy:
        log(f"queue occupancy: {occupancy['iterations']} iterations, ray queue {occupancy['ray_fill_mean']:.3f} "
            f"of {occupancy['capacity']} slots on average (min {occupancy['ray_fill_min']:.3f}), shadow queue "
            f"{occupancy['shadow_fill_mean']:.3f}; shares {occupancy['share']}")
    inst = r.RenderFrame(settings, view, ptcore.FRAME_INSTRUMENT)
    ext = inst["extension_rays"]
    nodes_per_ray = inst["closest_nodes"] / max(1, ext)
    tris_per_ray = inst["closest_tris"] / max(1, ext)
    bytes_per_ray = RAY_IN_B + HIT_OUT_B + node_b * nodes_per_ray + TRI_B * tris_per_ray
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
    # the roofline of the dominant traversal kernel (the larger share of the
    # profiled frame), and the other one beside it
    sh_launches = launches
    sh_bytes_per_ray = SHADOW_IN_B + node_b_any * sh_nodes + TRI_B * sh_tris
    sh_avg_ms = prof["shadow_ms"] / sh_launches
    sh_bytes_per_launch = sh_bytes_per_ray * prof["shadow_rays"] / sh_launches
    sh_achieved = sh_bytes_per_launch / (sh_avg_ms * 1e-3) / 1e9

    def roof(kind, kernel, achieved_gbs, bpr, nodes, tris, avg_ms, rays_per_launch, node_bytes):
        # the committed counter runs are of the production frame: no traffic figure for other scenes
        traffic = measured_traffic(kind, node_bytes) if a.config == "production" else None
        # the same launch's measured HBM bytes (rocprofv3 counters, per launch)
        # over its live duration: what HBM actually carried
        hbm = None
        if traffic:
            hbm = traffic["hbm_bytes_per_ray"] * rays_per_launch / (avg_ms * 1e-3) / 1e9
        return {"bound": "hbm", "achieved": round(achieved_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved_gbs / HBM_PEAK_GBS, 4),
                "hbm_achieved": round(hbm, 1) if hbm else None,
                "hbm_frac": round(hbm / HBM_PEAK_GBS, 4) if hbm else None,
                "traffic": traffic["hbm_bytes_per_launch"] if traffic else None, "traffic_detail": traffic,
                "kernel": kernel, "bytes_per_ray": round(bpr, 1), "node_bytes": node_bytes,
                "nodes_per_ray": round(nodes, 3), "tris_per_ray": round(tris, 3),
                "avg_launch_ms": round(avg_ms, 4), "rays_per_launch": round(rays_per_launch, 1)}

    r_closest = roof("closest", "k_trace<closest>", achieved, bytes_per_ray, nodes_per_ray, tris_per_ray, avg_launch_ms,
                     prof["extension_rays"] / launches, node_b)
    r_any = roof("any", "k_trace<any>", sh_achieved, sh_bytes_per_ray, sh_nodes, sh_tris, sh_avg_ms,
                 prof["shadow_rays"] / sh_launches, node_b_any)
    log(f"any-hit: avg {sh_avg_ms:.3f} ms per launch, {sh_bytes_per_ray:.0f} B/ray, {sh_achieved:.0f} GB/s algorithmic")
    if prof["shadow_ms"] >= prof["closest_ms"]:
        roofline, other = r_any, ("roofline_closest", r_closest)
    else:
        roofline, other = r_closest, ("roofline_any", r_any)

    serving = None
    if not a.no_serving and a.config == "production":
        # the metric's latency half: a compact serving run over two resident
        # config-1 rooms, requests up to 1920x1080 and 64 spp
        rooms, reqs = serving_workload([250_000, 1_000_000], 40, 4, resolutions=scenes.SERVE_RESOLUTIONS[:3],
                                       spp=scenes.SERVE_SPP[:5])
        rs = ptcore.Renderer()
        upload_all(rs, rooms)
        serving = serving_summary(rs, reqs, (0.7,), 21)
        rs.close()

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
        "roofline": roofline, other[0]: other[1], "cpu_baseline": cpu, "queue_occupancy": occupancy,
    }
    if serving:
        out["p50_ms"] = serving["rates"][0]["p50_ms"]
        out["p99_ms"] = serving["rates"][0]["p99_ms"]
        out["serving"] = dict(serving, workload="2 resident config-1 rooms (250k, 1M triangles), 40 Poisson "
                                                 "requests 512x512-1920x1080, 4-64 spp, depth 8")
    print(json.dumps(out), flush=True)
    r.close()


if __name__ == "__main__":
    main()
