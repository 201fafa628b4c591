"""Parity of the HIP core against the CPU oracle, through the C ABI."""
import math
import os

import numpy as np
import pytest

import ptcore
import scenes

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def c1(gpu_renderer, cornell):
    gpu_renderer.InitializeRenderingBuffers(64, 64)
    gpu_renderer.InitializeSceneBuffers(cornell)
    return gpu_renderer


def test_c1_closest_hit_bit_exact_1m_rays(c1, cornell_oracle):
    """1,048,576 seeded rays: (prim, t, u, v) equal the oracle's (t, prim) minimum bit for bit."""
    rays = scenes.random_rays(1 << 20, (-1, 0, -1), (1, 2, 1), 101)
    got = c1.trace_closest(rays)
    want = cornell_oracle.closest(rays, brute=True)
    assert np.array_equal(got, want)
    ref = cornell_oracle.closest(rays)            # reference traversal order
    diff = (got != ref).any(1)
    assert np.array_equal(got[diff, 1], ref[diff, 1])   # only exact ties in t differ


def test_c1_any_hit_matches(c1, cornell, cornell_oracle):
    rays = scenes.random_rays(1 << 18, (-1, 0, -1), (1, 2, 1), 102)
    rays[:, 7] = np.random.default_rng(1).uniform(0.01, 3.0, len(rays)).astype(np.float32)
    assert np.array_equal(c1.trace_any(rays), cornell_oracle.any(rays))
    sh = scenes.shadow_rays(cornell, 262144, 103)
    assert np.array_equal(c1.trace_any(sh), cornell_oracle.any(sh))


def test_c1_queue_membership(c1, cornell_oracle):
    rays = scenes.random_rays(1 << 18, (-1, 0, -1), (1, 2, 1), 104)
    q, counts = c1.classify(rays)
    want = cornell_oracle.classify(rays)
    assert np.array_equal(q, want)
    assert np.array_equal(counts, np.bincount(want, minlength=5))


def test_c1_alias_table_and_picks(c1, cornell_oracle):
    bins = c1.alias_table()
    want = cornell_oracle.alias_table()
    assert bins.tobytes() == want.tobytes()
    import oracle
    u = np.random.default_rng(5).random(1 << 20).astype(np.float32)
    u[:4] = [0.0, 0.5, 0.99999994, 0.25]
    idx, pmf = c1.alias_sample(u)
    oidx, opmf = oracle.alias_sample(want, u)
    assert np.array_equal(idx, oidx) and np.array_equal(pmf, opmf)


def test_golden_fixtures(c1):
    path = os.path.join(GOLDEN, "c1_hits_4096.npz")
    g = np.load(path)
    assert np.array_equal(c1.trace_closest(g["rays"]), g["hits"])
    assert np.array_equal(c1.trace_any(g["shadow_rays"]), g["occluded"])
    q, _ = c1.classify(g["rays"])
    assert np.array_equal(q, g["queue"])


def test_empty_batch_and_miss_rays(c1):
    assert c1.trace_closest(np.zeros((0, 8), np.float32)).shape == (0, 4)
    rays = scenes.random_rays(1000, (5, 5, 5), (6, 6, 6), 1)
    rays[:, 4:7] = np.abs(rays[:, 4:7]) + 0.1   # pointing away from the room
    h = c1.trace_closest(rays)
    assert (h[:, 0] == ptcore.INVALID).all()
    q, counts = c1.classify(rays)
    assert (q == ptcore.QUEUE_REGEN).all() and counts[4] == 1000


def _relmse(a, b, eps=1e-2):
    return float(np.mean((a - b) ** 2 / (b ** 2 + eps)))


def test_c1_render_matches_oracle(c1, cornell, cornell_oracle):
    st = dict(width=64, height=64, samples_per_pixel=256, path_depth=8, use_nee=1, rr_start_depth=3, seed=1)
    stats = c1.RenderFrame(st, cornell.view)
    hdr = c1.GetFrameHdr(64, 64).reshape(-1, 4)
    ref, ext, sh = cornell_oracle.render(st, cornell.view)
    # same sample streams and the same float operations: equal ray counts,
    # images equal up to the order of the float additions into a pixel
    assert stats["extension_rays"] == ext and stats["shadow_rays"] == sh
    rm = _relmse(hdr[:, :3], ref[:, :3])
    assert rm < 1e-10, rm
    mg, mo = hdr[:, :3].mean(), ref[:, :3].mean()
    se = math.sqrt(ref[:, :3].mean(1).var() / len(ref))
    assert abs(mg - mo) < 3 * se


@pytest.mark.parametrize("change, spp", [
    (dict(path_depth=1), 64),                           # camera hit and its NEE only
    (dict(path_depth=2, use_nee=0), 64),                # one bounce, emission found by BSDF sampling
    (dict(use_nee=0, rr_start_depth=64), 64),           # roulette never starts
    (dict(rr_start_depth=1), 64),                       # roulette from the first bounce
    (dict(samples_per_pixel=10), 9),                    # rounded down to a square (Renderer.cpp:366-367)
    (dict(width=40, height=16), 64),                    # another aspect ratio
])
def test_c1_render_settings_match_oracle(c1, cornell, cornell_oracle, change, spp):
    """Render settings at their edges: the same sample streams give the
    oracle's image and ray counts."""
    st = dict(width=32, height=24, samples_per_pixel=64, path_depth=8, use_nee=1, rr_start_depth=3, seed=5)
    st.update(change)
    w, h = st["width"], st["height"]
    stats = c1.RenderFrame(st, cornell.view)
    hdr = c1.GetFrameHdr(w, h).reshape(-1, 4)
    ref, ext, sh = cornell_oracle.render(st, cornell.view)
    assert stats["samples"] == w * h * spp
    assert stats["extension_rays"] == ext and stats["shadow_rays"] == sh
    rm = _relmse(hdr[:, :3], ref[:, :3])
    assert rm < 1e-10, rm


def test_c1_render_graph_equals_stage_by_stage(c1, cornell):
    st = dict(width=32, height=32, samples_per_pixel=16, path_depth=8, use_nee=1, seed=2)
    s1 = c1.RenderFrame(st, cornell.view)
    a = c1.GetFrameHdr(32, 32)
    s2 = c1.RenderFrame(st, cornell.view, ptcore.FRAME_NO_GRAPH)
    b = c1.GetFrameHdr(32, 32)
    s3 = c1.RenderFrame(st, cornell.view, ptcore.FRAME_PROFILE)
    c = c1.GetFrameHdr(32, 32)
    for s in (s2, s3):
        assert s["extension_rays"] == s1["extension_rays"] and s["shadow_rays"] == s1["shadow_rays"]
    assert np.allclose(a, b, rtol=1e-4, atol=1e-5) and np.allclose(a, c, rtol=1e-4, atol=1e-5)
    assert s1["samples"] == 32 * 32 * 16


def test_c1_small_pool_regeneration(c1, cornell):
    """A pool smaller than the sample count forces path regeneration."""
    st = dict(width=32, height=32, samples_per_pixel=16, path_depth=8, use_nee=1, seed=2)
    big = c1.RenderFrame(st, cornell.view)
    a = c1.GetFrameHdr(32, 32)
    tun = c1.get_tuning()
    try:
        c1.set_tuning(pool_size=1024)
        small = c1.RenderFrame(st, cornell.view)
        b = c1.GetFrameHdr(32, 32)
    finally:
        c1.set_tuning(**tun)
    assert small["samples"] == big["samples"] == 32 * 32 * 16
    assert small["extension_rays"] == big["extension_rays"]
    assert small["iterations"] > big["iterations"]
    assert np.allclose(a, b, rtol=1e-4, atol=1e-5)


def test_async_simulation_lifecycle(c1, cornell, tmp_path):
    st = dict(width=32, height=32, samples_per_pixel=4, path_depth=4, use_nee=1, seed=1)
    views = [cornell.view, cornell.view]
    c1.StartSimulation(st, cornell.view, views, str(tmp_path))
    with pytest.raises(ptcore.PtError) as e:
        c1.StartSimulation(st, cornell.view, views, str(tmp_path))
    assert e.value.status == 2 or not c1.IsRendering()
    import time
    t0 = time.time()
    while c1.IsRendering() and time.time() - t0 < 60:
        time.sleep(0.01)
    c1.StopRendering()
    p = c1.GetProgress()
    assert p["done_frames"] == 2 and p["total_frames"] == 2
    assert c1.PeekLastError() is None
    pngs = sorted(os.listdir(tmp_path))
    assert pngs == ["frame_0000.png", "frame_0001.png"]
    assert open(tmp_path / pngs[0], "rb").read(8) == b"\x89PNG\r\n\x1a\n"
    with pytest.raises(ptcore.PtError) as e:
        c1.ResumeSimulation(5)
    assert e.value.status == 3
    c1.ResumeSimulation(1)
    c1.StopRendering()
    frame = c1.GetFrame(32, 32)
    assert (frame >> 24 == 255).all()


def test_unsupported_material_is_rejected(gpu_renderer, cornell):
    s = scenes.cornell_box()
    mats = s.materials.copy()
    mats[0]["model"] = scenes.MODEL_PHONG
    s.materials = mats
    with pytest.raises(ptcore.PtError) as e:
        gpu_renderer.InitializeSceneBuffers(s)
    assert e.value.status == 6
    gpu_renderer.InitializeSceneBuffers(cornell)   # restore


def test_stack_overflow_is_reported(cornell, cornell_oracle, oracle_lib):
    """A traversal stack capped below what the BVH needs (pt_tuning.stack_limit)
    drops subtrees; every call that traverses then fails with "traversal stack
    overflow" instead of returning wrong hits, and the default cap serves again.
    The traversals of the fp32 BVH4 (the room, closest hit on 4-wide nodes) need
    more than one entry everywhere; those of the compressed 8-wide tree keep at
    most one entry per level (a node's pending inner children) and rarely more
    than one in all: theirs is shown on twelve spheres (245,762 triangles),
    where a few of 16,384 rays need up to four entries in the oracle's walks."""
    rays = scenes.random_rays(1 << 16, (-1, 0, -1), (1, 2, 1), 301)
    st = dict(width=32, height=32, samples_per_pixel=4, path_depth=8, use_nee=1, seed=2)
    with ptcore.Renderer() as r:
        r.set_tuning(bvh_quantized=6)         # closest hit on the fp32 BVH4
        r.InitializeRenderingBuffers(32, 32)
        r.InitializeSceneBuffers(cornell)
        tun = r.get_tuning()
        r.set_tuning(stack_limit=1)
        for call in (lambda: r.trace_closest(rays), lambda: r.classify(rays),
                     lambda: r.RenderFrame(st, cornell.view),
                     lambda: r.RenderFrame(st, cornell.view, ptcore.FRAME_NO_GRAPH)):
            with pytest.raises(ptcore.PtError) as e:
                call()
            assert e.value.status == 4 and "traversal stack overflow" in str(e.value)
        r.set_tuning(**tun)
        assert np.array_equal(r.trace_closest(rays), cornell_oracle.closest(rays, tprim=True))
        r.RenderFrame(st, cornell.view)

    deep = scenes.spheres_scene(count=12, subdivisions=5, lights=64)
    with ptcore.Renderer() as r2:
        r2.InitializeSceneBuffers(deep)
        lo, hi = deep.bounds()
        segs = scenes.random_rays(1 << 14, lo - 1, hi + 1, 302)
        q8, order8 = r2.bvh8_quantized()
        o = oracle_lib.OracleScene(deep)
        occ_want, _, _, ms_any = o.bvh8q_any(q8, order8, segs)
        hits_want, _, _, ms_closest = o.bvh8q_closest(q8, order8, segs)
        assert ms_any.max() >= 2 and ms_closest.max() >= 2
        tun2 = r2.get_tuning()
        r2.set_tuning(stack_limit=1)
        for call in (lambda: r2.trace_any(segs), lambda: r2.trace_closest(segs)):
            with pytest.raises(ptcore.PtError) as e:
                call()
            assert e.value.status == 4 and "traversal stack overflow" in str(e.value)
        r2.set_tuning(**tun2)
        assert np.array_equal(r2.trace_any(segs), occ_want)
        assert np.array_equal(r2.trace_closest(segs), hits_want)


@pytest.mark.parametrize("make", [scenes.cornell_box, scenes.specular_room])
def test_queue_occupancy_per_bounce_equals_the_oracle(gpu_renderer, oracle_lib, make):
    """Per-iteration queue sizes of a profiled frame (the reference's
    per-iteration counter readback, Renderer.cpp:587-600).  With every sample
    in the pool at once, iteration k holds bounce k of every path, so the
    extension-ray count, the four material queues, the missed rays and the
    shadow rays of each iteration equal the oracle's per-bounce counts over the
    logged paths of all 4,096 samples."""
    s = make()
    r = gpu_renderer
    r.InitializeSceneBuffers(s)
    st = dict(width=32, height=32, samples_per_pixel=4, path_depth=8, use_nee=1, rr_start_depth=3, seed=31)
    stats = r.RenderFrame(st, s.view, ptcore.FRAME_PROFILE)
    occ = r.queue_occupancy()
    assert occ["extension"].sum() == stats["extension_rays"] and occ["shadow"].sum() == stats["shadow_rays"]
    assert (occ["material"].sum(1) + occ["missed"] == occ["extension"]).all()
    assert (occ["next"][:-1] == occ["extension"][1:]).all() and occ["next"][-1] == 0
    o = oracle_lib.OracleScene(s)
    qmap = {scenes.MODEL_DIFFUSE: 0, scenes.MODEL_GGX: 1, scenes.MODEL_MIRROR: 2, scenes.MODEL_EMISSIVE: 3}
    want = np.zeros((len(occ), 6), np.int64)     # 4 material queues, missed, shadow
    for sample in range(32 * 32 * 4):
        rec, _ = o.trace_path(st, s.view, sample)
        for k in range(len(rec)):
            d = rec["depth"][k]
            if rec["kind"][k] == 1:
                want[d, 5] += 1
                continue
            prim = rec["result"][k].view(np.uint32)[0]
            if prim == ptcore.INVALID:
                want[d, 4] += 1
            else:
                want[d, qmap[int(s.materials["model"][s.material_ids[prim]])]] += 1
    got = np.concatenate([occ["material"], occ["missed"][:, None], occ["shadow"][:, None]], 1)
    assert np.array_equal(got, want), (got, want)
    assert occ["extension"][0] == 32 * 32 * 4 and (occ["capacity"] >= 4096).all()


@pytest.fixture(scope="module")
def c2_pair(oracle_lib):
    s = scenes.production_scene()
    r = ptcore.Renderer()
    r.InitializeRenderingBuffers(64, 64)
    r.InitializeSceneBuffers(s)
    o = oracle_lib.OracleScene(s)
    yield s, r, o
    r.close()


def test_c2_closest_hit_16m_rays(c2_pair):
    """Production scene (983,042 triangles), 16,777,216 seeded rays: every hit
    (prim, t, u, v) equals the oracle's (t, prim id) minimum bit for bit; that
    minimum equals brute force on a subset, and the reference's traversal order
    differs from it only at exact ties in t.  Every hit is also re-derived with
    the same float32 arithmetic in numpy."""
    s, r, o = c2_pair
    rays = scenes.random_rays(1 << 24, (-10, -10, -10), (110, 110, 110), 201)
    hits = r.trace_closest(rays)
    want = o.closest(rays, tprim=True)
    diff = (hits != want).any(1)
    assert not diff.any(), (int(diff.sum()), hits[diff][:4], want[diff][:4])
    sub = np.random.default_rng(0).choice(len(rays), 1024, replace=False)
    assert np.array_equal(o.closest(rays[sub], brute=True), want[sub])
    ref = o.closest(rays)
    tie = (ref != want).any(1)
    assert np.array_equal(ref[tie, 1], want[tie, 1]) and tie.sum() < 100
    prim, t, u, v = ptcore.hits_as_fields(hits)
    hit = prim != ptcore.INVALID
    assert hit.mean() > 0.01
    # recompute Moller-Trumbore in float32 for the reported triangle (no contraction in numpy)
    R = rays[hit]
    P = s.positions[prim[hit]]
    o3, d = R[:, 0:3], R[:, 4:7]
    v0 = P[:, 0]
    e1 = P[:, 1] - P[:, 0]
    e2 = P[:, 2] - P[:, 0]

    def cross(a, b):
        return np.stack([a[:, 1] * b[:, 2] - a[:, 2] * b[:, 1], a[:, 2] * b[:, 0] - a[:, 0] * b[:, 2],
                         a[:, 0] * b[:, 1] - a[:, 1] * b[:, 0]], 1)

    def dot(a, b):
        return (a[:, 0] * b[:, 0] + a[:, 1] * b[:, 1]) + a[:, 2] * b[:, 2]

    pv = cross(d, e2)
    inv = np.float32(1.0) / dot(e1, pv)
    sv = o3 - v0
    qv = cross(sv, e1)
    assert np.array_equal(dot(sv, pv) * inv, u[hit])
    assert np.array_equal(dot(d, qv) * inv, v[hit])
    assert np.array_equal(dot(e2, qv) * inv, t[hit])


def test_c2_any_hit_matches_oracle(c2_pair):
    """Any hit on the production scene: 1,048,576 shadow segments from surface
    points to points on the 1,024 emitters, and 2^20 random rays with random
    t_max, occlusion equal to the oracle's for every ray (Intersect.cu:125-179)."""
    s, r, o = c2_pair
    sh = scenes.shadow_rays(s, 1 << 20, 203)
    occ = r.trace_any(sh)
    assert np.array_equal(occ, o.any(sh))
    assert 0.05 < occ.mean() < 0.95
    rays = scenes.random_rays(1 << 20, (-10, -10, -10), (110, 110, 110), 204)
    rays[:, 7] = np.random.default_rng(2).uniform(0.1, 150.0, len(rays)).astype(np.float32)
    assert np.array_equal(r.trace_any(rays), o.any(rays))


def _closest_hit_walk(r, o, rays):
    """The oracle's walk of the closest-hit tree a renderer traverses: the 8-wide
    compressed nodes where pt_tuning.bvh_quantized has bit 3 (the default) and
    they were built, else the 128-byte BVH4 nodes."""
    if r.get_tuning()["bvh_quantized"] & 8 and r.bvh_info()["bvh8_node_count"]:
        return o.bvh8q_closest(*r.bvh8_quantized(), rays)
    nodes, order = r.bvh4()
    return o.bvh4_closest(nodes, order, rays)


def _any_hit_walk(r, o, rays):
    """The oracle's walk of the any-hit tree a renderer traverses: the 8-wide
    quantized nodes where they were built, else the 4-wide quantized ones."""
    if r.bvh_info()["bvh8_node_count"]:
        occ, nv, nt, _ = o.bvh8q_any(*r.bvh8_quantized(), rays)
    else:
        _, order = r.bvh4()
        occ, nv, nt = o.bvh4q_any(r.bvh4_quantized(), order, rays)
    return occ, nv, nt


def test_c2_visit_counts_equal_the_oracle_walk(c2_pair):
    """The node and triangle visits behind the roofline's algorithmic bytes:
    per ray, the kernels' COUNT counters equal the oracle's walk of the same
    uploaded trees in the same order, for any hit over the default quantized
    nodes (262,144 shadow segments to the emitters) and closest hit over the
    default closest-hit tree (262,144 random rays); the results equal brute force
    on a subset."""
    s, r, o = c2_pair
    sh = scenes.shadow_rays(s, 1 << 18, 205)
    occ, nv, nt = r.trace_any_counts(sh)
    oocc, onv, ont = _any_hit_walk(r, o, sh)
    assert np.array_equal(occ, oocc) and np.array_equal(nv, onv) and np.array_equal(nt, ont)
    assert np.array_equal(occ[:512], o.any(sh[:512], brute=True))
    rays = scenes.random_rays(1 << 18, (-10, -10, -10), (110, 110, 110), 206)
    hits, cnv, cnt = r.trace_closest_counts(rays)
    ohits, ocnv, ocnt, _ = _closest_hit_walk(r, o, rays)
    assert np.array_equal(hits, ohits) and np.array_equal(cnv, ocnv) and np.array_equal(cnt, ocnt)
    assert np.array_equal(hits[:512], o.closest(rays[:512], brute=True))
    print(f"any hit: {nv.mean():.3f} nodes, {nt.mean():.3f} triangles per shadow segment; "
          f"closest hit: {cnv.mean():.3f} nodes, {cnt.mean():.3f} triangles per random ray")


def test_c2_closest_hit_on_the_8wide_tree_equals_the_oracle_walk(c2_pair):
    """Closest hit over the compressed 8-wide nodes (pt_tuning.bvh_quantized bit
    3, the default): per ray, (prim, t, u, v) and node / triangle visits equal
    the oracle's walk of the uploaded tree on 262,144 random rays (the walk itself
    equals brute force), and the hits equal the fp32 BVH4 kernel's on 2^20 rays."""
    s, _, o = c2_pair
    with ptcore.Renderer() as r, ptcore.Renderer() as r4:
        r4.set_tuning(bvh_quantized=6)
        r4.InitializeRenderingBuffers(64, 64)
        r4.InitializeSceneBuffers(s)
        r.set_tuning(bvh_quantized=14)
        r.InitializeRenderingBuffers(64, 64)
        r.InitializeSceneBuffers(s)
        q8, order8 = r.bvh8_quantized()
        rays = scenes.random_rays(1 << 18, (-10, -10, -10), (110, 110, 110), 207)
        hits, nv, nt = r.trace_closest_counts(rays)
        ohits, onv, ont, ms = o.bvh8q_closest(q8, order8, rays)
        assert np.array_equal(hits, ohits) and np.array_equal(nv, onv) and np.array_equal(nt, ont)
        assert np.array_equal(ohits[:512], o.closest(rays[:512], brute=True))
        big = scenes.random_rays(1 << 20, (-10, -10, -10), (110, 110, 110), 208)
        assert np.array_equal(r.trace_closest(big), r4.trace_closest(big))
        print(f"closest hit on the 8-wide tree: {nv.mean():.3f} nodes, {nt.mean():.3f} triangles per random ray")


@pytest.mark.parametrize("quantized", [2, 6])
def test_c2_any_hit_node_formats_equal_the_oracle_walk(c2_pair, quantized):
    """Any hit over the 64-byte 4-wide quantized nodes (tuning bit 1) and over
    the 128-byte 8-wide ones (bit 2): per ray, occlusion and node / triangle
    visits equal the oracle's walk of the uploaded tree on 262,144 shadow
    segments, occlusion equals the oracle's on 2^20, and the 8-wide walk costs
    fewer node visits than the 4-wide one."""
    s, _, o = c2_pair
    with ptcore.Renderer() as r:
        r.set_tuning(bvh_quantized=quantized)
        r.InitializeRenderingBuffers(64, 64)
        r.InitializeSceneBuffers(s)
        info = r.bvh_info()
        _, order = r.bvh4()
        q4 = r.bvh4_quantized()
        sh = scenes.shadow_rays(s, 1 << 18, 205)
        occ, nv, nt = r.trace_any_counts(sh)
        if quantized & 4:
            q8, order8 = r.bvh8_quantized()
            assert info["bvh8_node_count"] == len(q8) and 1 <= info["bvh8_depth"] <= info["max_depth"]
            assert sorted(order8.tolist()) == list(range(s.triangle_count))
            oocc, onv, ont, ms = o.bvh8q_any(q8, order8, sh)
            assert ms.max() <= info["bvh8_depth"]
            _, nv4, _ = o.bvh4q_any(q4, order, sh)
            assert nv.mean() < nv4.mean()
        else:
            assert info["bvh8_node_count"] == 0
            oocc, onv, ont = o.bvh4q_any(q4, order, sh)
        assert np.array_equal(occ, oocc) and np.array_equal(nv, onv) and np.array_equal(nt, ont)
        big = scenes.shadow_rays(s, 1 << 20, 203)
        assert np.array_equal(r.trace_any(big), o.any(big))
        print(f"bvh_quantized={quantized}: {nv.mean():.3f} nodes, {nt.mean():.3f} triangles per shadow segment")


def test_c2_queue_membership_and_alias(c2_pair, oracle_lib):
    """Queue membership of 2^20 rays equal to the oracle's for every ray; alias
    bins byte-equal and 2^20 picks equal."""
    s, r, o = c2_pair
    rays = scenes.random_rays(1 << 20, (-10, -10, -10), (110, 110, 110), 202)
    q, counts = r.classify(rays)
    want = o.classify(rays)
    assert np.array_equal(q, want)
    assert np.array_equal(counts, np.bincount(want, minlength=5))
    assert (np.bincount(want, minlength=5)[:4] > 20).all()
    bins = r.alias_table()
    assert bins.tobytes() == o.alias_table().tobytes()
    u = np.random.default_rng(9).random(1 << 20).astype(np.float32)
    idx, pmf = r.alias_sample(u)
    oi, op = oracle_lib.alias_sample(bins, u)
    assert np.array_equal(idx, oi) and np.array_equal(pmf, op)


def _exact_occlusion(scene, ray):
    """Whether the segment [t_min, t_max) of `ray` meets any triangle, solved in
    float64 (Cramer's rule over every triangle)."""
    P = scene.positions.astype(np.float64)
    o, d = ray[0:3].astype(np.float64), ray[4:7].astype(np.float64)
    v0 = P[:, 0]
    e1, e2 = P[:, 1] - v0, P[:, 2] - v0
    pv = np.cross(d[None], e2)
    det = (e1 * pv).sum(1)
    ok = np.abs(det) > 1e-300
    inv = np.where(ok, 1.0 / np.where(ok, det, 1.0), 0.0)
    sv = o[None] - v0
    u = (sv * pv).sum(1) * inv
    q = np.cross(sv, e1)
    v = (q * d[None]).sum(1) * inv
    t = (e2 * q).sum(1) * inv
    hit = ok & (u >= 0) & (v >= 0) & (u + v <= 1) & (t >= ray[3]) & (t < ray[7])
    return bool(hit.any())


def _explain_mismatch(scene, r, o, st, pixel):
    """Replays every ray of the pixel's samples (oracle path log) through the
    core's hooks: extension rays must hit identically; a shadow ray may differ
    only where float32 Moller-Trumbore over all triangles disagrees with the
    exact float64 answer (a false hit or miss of the float test on a triangle's
    edge, which box culling then keeps or drops depending on the tree)."""
    explained = 0
    for c in range(st["samples_per_pixel"]):
        rec, _ = o.trace_path(st, scene.view, pixel * st["samples_per_pixel"] + c)
        ext, sh = rec["ray"][rec["kind"] == 0], rec["ray"][rec["kind"] == 1]
        if len(ext):
            assert np.array_equal(r.trace_closest(ext), rec["result"][rec["kind"] == 0].view(np.uint32))
        if len(sh):
            want = (rec["result"][rec["kind"] == 1][:, 0] > 0).astype(np.uint32)
            for k in np.nonzero(r.trace_any(sh) != want)[0]:
                brute32 = bool(o.any(sh[k:k + 1], brute=True)[0])
                assert brute32 != _exact_occlusion(scene, sh[k]), (pixel, c, sh[k].tolist())
                explained += 1
    return explained


def test_c2_production_frame_matches_oracle(c2_pair):
    """The benchmarked frame (1920 x 1080, 64 spp, depth 8, NEE, roulette from
    bounce 3): a band of 16 rows through the image centre (30,720 pixels,
    1,966,080 samples) against the oracle pixel by pixel, and the same view at
    240 x 135 whole with equal ray counts.  Pixels agree to the rounding of
    their float sums except where a shadow ray's float32 triangle test is
    itself wrong (edge-grazing false hits, decided by box culling): at most
    1 in 10,000 pixels, each one traced to such a ray."""
    s, r, o = c2_pair
    st = dict(s.settings)
    r.InitializeRenderingBuffers(1920, 1080)
    r.RenderFrame(st, s.view)
    hdr = r.GetFrameHdr(1920, 1080).reshape(-1, 4)
    lo, hi = 1920 * 532, 1920 * 548
    ref, _, _ = o.render(st, s.view, pixel_begin=lo, pixel_end=hi, tprim=True)
    g = hdr[lo:hi, :3]
    assert (ref[:, :3] > 0).mean() > 0.5
    bad = np.nonzero(~np.isclose(g, ref[:, :3], rtol=2e-5, atol=1e-6).all(1))[0]
    assert len(bad) <= 1e-4 * len(g), len(bad)
    for k in bad:
        assert _explain_mismatch(s, r, o, st, lo + int(k)) > 0
    assert _relmse(g, ref[:, :3]) < 1e-8
    se = math.sqrt(ref[:, :3].mean(1).var() / len(ref))
    assert abs(g.mean() - ref[:, :3].mean()) < 3 * se
    small = dict(st, width=240, height=135)
    stats = r.RenderFrame(small, s.view)
    hdr = r.GetFrameHdr(240, 135).reshape(-1, 4)
    ref, ext, sh = o.render(small, s.view, tprim=True)
    assert stats["extension_rays"] == ext and stats["shadow_rays"] == sh
    bad = np.nonzero(~np.isclose(hdr[:, :3], ref[:, :3], rtol=2e-5, atol=1e-6).all(1))[0]
    assert len(bad) <= 1e-4 * len(hdr), len(bad)
    for k in bad:
        assert _explain_mismatch(s, r, o, small, int(k)) > 0
