"""Serving loop: continuous batching of render requests over resident scenes.

Each request's result must equal the same view rendered alone (only the order
of float additions into a pixel differs), requests that arrive while others
run must overlap on the device, and admission must respect the memory
budget."""
import time

import numpy as np
import pytest

import ptcore
import scenes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def spheres():
    return scenes.spheres_scene(count=6, subdivisions=2, lights=64, seed=7)


@pytest.fixture(scope="module")
def server(cornell, spheres):
    r = ptcore.Renderer()
    r.InitializeRenderingBuffers(32, 32)
    r.InitializeSceneBuffers(cornell)
    sid = r.AddScene(spheres)
    yield r, sid
    try:
        r.ServerStop()
    except ptcore.PtError:
        pass
    r.close()


def _requests(cornell, spheres, sid):
    side = scenes.View(cornell.view.fov_y, (0.6, 1.2, 3.0),
                       scenes.look_at_quaternion((0.6, 1.2, 3.0), (0.0, 0.8, 0.0)))
    return [
        (0, dict(width=48, height=40, samples_per_pixel=64, path_depth=8, use_nee=1, seed=3), cornell.view),
        (sid, dict(width=40, height=30, samples_per_pixel=16, path_depth=6, use_nee=1, seed=5), spheres.view),
        (0, dict(width=24, height=16, samples_per_pixel=9, path_depth=5, use_nee=0, rr_start_depth=2, seed=4), side),
        (sid, dict(width=8, height=8, samples_per_pixel=1, path_depth=3, use_nee=0, seed=6), spheres.view),
        (0, dict(width=64, height=64, samples_per_pixel=16, path_depth=4, use_nee=1, seed=8), side),
        (sid, dict(width=32, height=32, samples_per_pixel=25, path_depth=8, use_nee=1, seed=9), spheres.view),
    ]


def _alone(r, items):
    out = []
    for it in items:
        (hdr,), _, (ldr,) = r.RenderBatch([it], rgba=True)
        out.append((hdr, ldr))
    return out


def _check_equal(got_hdr, got_ldr, want_hdr, want_ldr):
    assert got_hdr.shape == want_hdr.shape
    assert np.allclose(got_hdr, want_hdr, rtol=1e-4, atol=1e-5)
    # RGBA8 of nearly equal radiance: bytes differ by at most one step
    for sh in (0, 8, 16):
        a = (got_ldr >> sh) & 0xFF
        b = (want_ldr >> sh) & 0xFF
        assert np.abs(a.astype(np.int64) - b.astype(np.int64)).max() <= 1
    assert ((got_ldr >> 24) == 255).all()


def test_requests_equal_views_alone_under_chunking(server, cornell, spheres):
    """Tiny chunks and window force every request to be split and
    interleaved with the others; arrivals are staggered."""
    r, sid = server
    items = _requests(cornell, spheres, sid)
    want = _alone(r, items)
    r.ServerStart(chunk_samples=3001, window_samples=20000)
    ids = []
    for k, it in enumerate(items):
        ids.append(r.ServerSubmit(*it, hdr=True, rgba=True))
        if k == 2:
            time.sleep(0.05)
    infos = [r.ServerWait(i) for i in ids]
    stats = r.ServerStop()
    for rid, info, (wh, wl) in zip(ids, infos, want):
        assert info["state"] == ptcore.REQUEST_DONE, info
        hdr, ldr = r.ServerResult(rid)
        _check_equal(hdr, ldr, wh, wl)
        assert info["submit_ms"] <= info["admit_ms"] <= info["done_ms"]
    assert stats["requests_done"] == len(items) and stats["requests_failed"] == 0
    assert stats["samples"] == sum(it[1]["width"] * it[1]["height"] * it[1]["samples_per_pixel"] for it in items)
    assert stats["extension_rays"] > stats["samples"] // 2


def test_requests_overlap_on_the_device(server, cornell, spheres):
    r, sid = server
    items = _requests(cornell, spheres, sid)
    r.ServerStart()
    ids = [r.ServerSubmit(*it, hdr=False) for it in items]
    infos = [r.ServerWait(i) for i in ids]
    r.ServerStop()
    assert all(i["state"] == ptcore.REQUEST_DONE for i in infos)
    first_done = min(i["done_ms"] for i in infos)
    # continuous batching: several requests were admitted before any finished
    assert sum(i["admit_ms"] < first_done for i in infos) >= 3
    # the smallest request is served before the largest one, though it was submitted later
    small = infos[3]
    big = infos[0]
    assert small["done_ms"] <= big["done_ms"]


def test_admission_respects_memory_budget(server, cornell, spheres):
    r, sid = server
    items = _requests(cornell, spheres, sid)
    largest = max(it[1]["width"] * it[1]["height"] for it in items)
    r.ServerStart(accum_budget_bytes=20 * largest)
    ids = [r.ServerSubmit(*it) for it in items]
    infos = [r.ServerWait(i) for i in ids]
    r.ServerStop()
    assert all(i["state"] == ptcore.REQUEST_DONE for i in infos)
    # the budget holds the largest request: two requests whose pixels do not
    # fit together never run at once
    px = [it[1]["width"] * it[1]["height"] for it in items]
    for a in range(len(items)):
        for b in range(len(items)):
            if a != b and px[a] + px[b] > largest:
                ia, ib = infos[a], infos[b]
                overlap = ia["admit_ms"] < ib["done_ms"] and ib["admit_ms"] < ia["done_ms"]
                assert not overlap, (a, b)


def test_server_errors_and_lifecycle(server, cornell, spheres):
    r, sid = server
    st = dict(width=16, height=16, samples_per_pixel=4, path_depth=4, use_nee=1, seed=1)
    with pytest.raises(ptcore.PtError) as e:
        r.ServerSubmit(0, st, cornell.view)
    assert e.value.status == 2
    r.ServerStart(accum_budget_bytes=20 * 16 * 16)
    try:
        with pytest.raises(ptcore.PtError) as e:
            r.ServerStart()
        assert e.value.status == 2
        with pytest.raises(ptcore.PtError) as e:
            r.ServerSubmit(33, st, cornell.view)
        assert e.value.status == 1
        with pytest.raises(ptcore.PtError) as e:
            r.ServerSubmit(0, dict(st, width=17), cornell.view)   # larger than the budget
        assert e.value.status == 1
        with pytest.raises(ptcore.PtError) as e:
            r.RenderFrame(st, cornell.view)
        assert e.value.status == 2
        with pytest.raises(ptcore.PtError) as e:
            r.StopRendering()
        assert e.value.status == 2
        rid = r.ServerSubmit(0, st, cornell.view)
        info = r.ServerWait(rid)
        assert info["state"] == ptcore.REQUEST_DONE
        s = r.ServerStats()
        assert s["requests_done"] == 1 and s["busy_ms"] <= s["wall_ms"]
    finally:
        r.ServerStop()
    with pytest.raises(ptcore.PtError) as e:
        r.ServerStop()
    assert e.value.status == 2
    # the renderer works normally afterwards
    (hdr,), _ = r.RenderBatch([(sid, st, spheres.view)])
    assert np.isfinite(hdr).all()


def test_served_request_equals_the_oracle(server, cornell, oracle_lib):
    """A request that went through the server (split into chunks and mixed with
    another) against the oracle's render of the same view (Renderer.cpp:330-399
    renders one request at a time): equal to the rounding of the float sums."""
    r, sid = server
    st = dict(width=48, height=40, samples_per_pixel=16, path_depth=8, use_nee=1, rr_start_depth=3, seed=21)
    r.ServerStart(chunk_samples=4001, window_samples=16000)
    rid = r.ServerSubmit(0, st, cornell.view, hdr=True)
    other = r.ServerSubmit(sid, dict(st, seed=22), cornell.view, hdr=True)
    info = r.ServerWait(rid)
    r.ServerWait(other)
    r.ServerStop()
    assert info["state"] == ptcore.REQUEST_DONE
    hdr, _ = r.ServerResult(rid)
    r.ServerResult(other)
    ref, _, _ = oracle_lib.OracleScene(cornell).render(st, cornell.view, tprim=True)
    assert np.allclose(hdr.reshape(-1, 4)[:, :3], ref[:, :3], rtol=2e-5, atol=1e-6)
