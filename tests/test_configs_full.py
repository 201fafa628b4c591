"""The batching, serving and guiding configurations at their configured sizes
(config 3: 32 scenes of 10k-2M triangles and 84 views; config 4: 8 resident
rooms of 1-20M triangles; config 5: a 200k-triangle room at 1280 x 720),
checked by properties that do not depend on the size: each view or request
equals the same view rendered alone, one of each against the oracle, and the
guided estimator's mean against plain BSDF sampling's."""
import gc
import math

import numpy as np
import pytest

import ptcore
import scenes

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.max(np.abs(a - b) / (np.abs(b) + 1e-3)))


def test_batch_workload_views_equal_views_alone(oracle_lib):
    """All 84 views of the 32 scenes in one wavefront; five of them (the
    smallest and largest scenes, every resolution class) equal to the view
    rendered alone, and a view of the smallest scene equal to the oracle."""
    sc, views = scenes.batch_workload()
    r = ptcore.Renderer()
    try:
        r.set_tuning(bvh_builder=ptcore.BVH_BUILDER_GPU)
        ids = [0]
        r.InitializeSceneBuffers(sc[0])
        for s in sc[1:]:
            ids.append(r.AddScene(s))
        items = [(ids[i], st, v) for i, st, v in views]
        tri = np.array([sc[i].triangle_count for i, _, _ in views])
        area = np.array([st["width"] * st["height"] for _, st, _ in views])
        pick = sorted({int(np.argmin(tri)), int(np.argmax(tri)), int(np.argmax(area)), int(np.argmin(area)),
                       len(views) // 2})
        batch, bst = r.RenderBatch(items, hdr=[k in pick for k in range(len(items))])
        assert bst["samples"] == sum(st["width"] * st["height"] * (int(math.isqrt(st["samples_per_pixel"])) ** 2)
                                     for _, st, _ in views)
        for k in pick:
            (alone,), _ = r.RenderBatch([items[k]])
            assert np.allclose(batch[k], alone, rtol=1e-4, atol=1e-5), (k, _rel(batch[k], alone))
        k = int(np.argmin(tri))
        i, st, v = views[k]
        small = dict(st, width=64, height=64, samples_per_pixel=4)
        (got,), gst = r.RenderBatch([(ids[i], small, v)])
        ref, ext, sh = oracle_lib.OracleScene(sc[i]).render(small, v, tprim=True)
        assert gst["extension_rays"] == ext and gst["shadow_rays"] == sh
        assert np.allclose(got.reshape(-1, 4)[:, :3], ref[:, :3], rtol=2e-5, atol=1e-6)
    finally:
        r.close()


def test_serving_resident_rooms_of_1_to_20m_triangles(oracle_lib):
    """Eight rooms of 1.0-20.0M triangles resident at once (built on the GPU);
    requests to each room served concurrently equal the same views alone, and
    a request to the 1M-triangle room equals the oracle's render."""
    sizes = [int(x) for x in np.exp(np.linspace(np.log(1e6), np.log(20e6), 8))]
    r = ptcore.Renderer()
    try:
        r.set_tuning(bvh_builder=ptcore.BVH_BUILDER_GPU)
        ids, rooms_small = [], None
        for k, t in enumerate(sizes):
            room = scenes.cornell_scaled(t)
            if k == 0:
                r.InitializeSceneBuffers(room)
                ids.append(0)
                rooms_small = room
            else:
                ids.append(r.AddScene(room))
                del room
            gc.collect()
        assert rooms_small.triangle_count >= 900_000
        rng = np.random.default_rng(12)
        reqs = []
        for k, sid in enumerate(ids):
            req = scenes.serve_request(rng, 1, resolutions=((128, 128), (160, 90)), spp=(4, 9, 16))
            _, st, v = req
            reqs.append((sid, dict(st, seed=100 + k), v))
        want = [r.RenderBatch([q])[0][0] for q in reqs]
        r.ServerStart(chunk_samples=20011, window_samples=400000)
        rids = [r.ServerSubmit(*q, hdr=True) for q in reqs]
        infos = [r.ServerWait(i) for i in rids]
        stats = r.ServerStop()
        assert stats["requests_done"] == len(reqs) and stats["requests_failed"] == 0
        for rid, info, w in zip(rids, infos, want):
            assert info["state"] == ptcore.REQUEST_DONE
            hdr, _ = r.ServerResult(rid)
            assert np.allclose(hdr, w, rtol=1e-4, atol=1e-5), _rel(hdr, w)
        # the served request to the 1M-triangle room against the oracle
        sid, st, v = reqs[0]
        r.ServerStart()
        rid = r.ServerSubmit(sid, st, v, hdr=True)
        assert r.ServerWait(rid)["state"] == ptcore.REQUEST_DONE
        r.ServerStop()
        hdr, _ = r.ServerResult(rid)
        ref, _, _ = oracle_lib.OracleScene(rooms_small).render(st, v, tprim=True)
        assert np.allclose(hdr.reshape(-1, 4)[:, :3], ref[:, :3], rtol=2e-5, atol=1e-6)
    finally:
        r.close()


def test_guiding_room_full_size_is_unbiased():
    """The 200k-triangle guiding room at 1280 x 720 without NEE (nothing
    clamped): the online-guided render's image mean equals plain BSDF
    sampling's within 3 standard errors of their per-pixel difference."""
    room = scenes.guiding_room(200_000)
    assert room.triangle_count >= 190_000
    r = ptcore.Renderer()
    try:
        r.InitializeRenderingBuffers(1280, 720)
        r.InitializeSceneBuffers(room)
        st = dict(width=1280, height=720, samples_per_pixel=64, path_depth=8, use_nee=0, rr_start_depth=3, seed=7)
        r.RenderGuided(st, room.view, passes=4, guide_fraction=0.0)
        plain = r.GetFrameHdr(1280, 720)[..., :3].copy()
        gst, gs = r.RenderGuided(dict(st, seed=8), room.view, passes=4, train_steps=256, batch=16384,
                                 learning_rate=3e-3, guide_fraction=0.5, reset=True, seed=1)
        guided = r.GetFrameHdr(1280, 720)[..., :3].copy()
        assert gst["samples"] == 1280 * 720 * 64 and gs["records"] > 0 and math.isfinite(gs["loss_last"])
        assert np.isfinite(guided).all() and (guided >= 0).all()
        d = (guided - plain).mean(-1).ravel()
        assert abs(d.mean()) < 3 * d.std() / math.sqrt(d.size), (d.mean(), d.std())
        assert plain.mean() > 0
    finally:
        r.close()
