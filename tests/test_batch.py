"""Batched wavefront: many views of several resident scenes in one launch
sequence.  Each view's image must equal the image it gets rendered alone (the
sample streams are keyed by view seed and local sample, so only the order of
the float additions into a pixel differs), and a view of a non-default scene
must match the oracle."""
import math

import numpy as np
import pytest

import ptcore
import scenes

pytestmark = pytest.mark.gpu


def _relmse(a, b):
    return float(np.mean((a - b) ** 2 / (b ** 2 + 1e-2)))


@pytest.fixture(scope="module")
def spheres():
    return scenes.spheres_scene(count=6, subdivisions=2, lights=64, seed=7)


@pytest.fixture(scope="module")
def two_scenes(cornell, spheres):
    r = ptcore.Renderer()
    r.InitializeRenderingBuffers(32, 32)
    r.InitializeSceneBuffers(cornell)
    sid = r.AddScene(spheres)
    yield r, sid
    r.close()


def _items(cornell, spheres, sid):
    side = scenes.View(cornell.view.fov_y, (0.6, 1.2, 3.0),
                       scenes.look_at_quaternion((0.6, 1.2, 3.0), (0.0, 0.8, 0.0)))
    return [
        (0, dict(width=32, height=32, samples_per_pixel=16, path_depth=8, use_nee=1, seed=3), cornell.view),
        (sid, dict(width=40, height=30, samples_per_pixel=4, path_depth=6, use_nee=1, seed=5), spheres.view),
        (0, dict(width=24, height=16, samples_per_pixel=9, path_depth=5, use_nee=0, rr_start_depth=2, seed=4), side),
        (sid, dict(width=8, height=8, samples_per_pixel=1, path_depth=3, use_nee=0, seed=6), spheres.view),
    ]


def test_batch_views_equal_views_alone(two_scenes, cornell, spheres):
    r, sid = two_scenes
    items = _items(cornell, spheres, sid)
    batch, bst = r.RenderBatch(items)
    ext = 0
    for i, it in enumerate(items):
        alone, st = r.RenderBatch([it])
        ext += st["extension_rays"]
        assert alone[0].shape == batch[i].shape
        assert np.allclose(batch[i], alone[0], rtol=1e-4, atol=1e-5), i
        assert np.isfinite(batch[i]).all() and (batch[i][..., :3] >= 0).all()
    # path decisions depend only on (view seed, local sample): same ray counts
    assert bst["extension_rays"] == ext
    want = sum(it[1]["width"] * it[1]["height"] * it[1]["samples_per_pixel"] for it in items)
    assert bst["samples"] == want


def test_batch_graph_equals_stage_by_stage(two_scenes, cornell, spheres):
    r, sid = two_scenes
    items = _items(cornell, spheres, sid)[:3]
    a, sa = r.RenderBatch(items)
    b, sb = r.RenderBatch(items, ptcore.FRAME_NO_GRAPH)
    c, sc = r.RenderBatch(items, ptcore.FRAME_PROFILE)
    assert sa["extension_rays"] == sb["extension_rays"] == sc["extension_rays"]
    for x, y, z in zip(a, b, c):
        assert np.allclose(x, y, rtol=1e-4, atol=1e-5) and np.allclose(x, z, rtol=1e-4, atol=1e-5)


def test_batch_single_view_equals_render_frame(two_scenes, cornell):
    r, _ = two_scenes
    st = dict(width=32, height=32, samples_per_pixel=16, path_depth=8, use_nee=1, seed=11)
    r.RenderFrame(st, cornell.view)
    frame = r.GetFrameHdr(32, 32)
    (hdr,), _ = r.RenderBatch([(0, st, cornell.view)])
    assert np.allclose(hdr, frame, rtol=1e-4, atol=1e-5)
    # the batch overwrote the accumulation buffer of the single frame
    with pytest.raises(ptcore.PtError) as e:
        r.GetFrameHdr(32, 32)
    assert e.value.status == 2


def test_batch_second_scene_matches_oracle(two_scenes, spheres, oracle_lib):
    r, sid = two_scenes
    st = dict(width=24, height=24, samples_per_pixel=64, path_depth=6, use_nee=1, seed=9)
    hdr, stats = r.RenderBatch([(0, st, scenes.cornell_box().view), (sid, st, spheres.view)])
    ref, ext, _ = oracle_lib.OracleScene(spheres).render(st, spheres.view)
    got = hdr[1].reshape(-1, 4)
    rm = _relmse(got[:, :3], ref[:, :3])
    assert rm < 1e-3, rm
    mg, mo = got[:, :3].mean(), ref[:, :3].mean()
    se = math.sqrt(ref[:, :3].mean(1).var() / len(ref))
    assert abs(mg - mo) < 3 * se + 1e-6


def test_batch_rgba_and_errors(two_scenes, cornell, spheres):
    r, sid = two_scenes
    items = _items(cornell, spheres, sid)[:2]
    hdr, _, ldr = r.RenderBatch(items, rgba=True)
    for h, l in zip(hdr, ldr):
        assert l.shape == h.shape[:2] and (l >> 24 == 255).all()
        # the red byte is a non-decreasing function of the red radiance
        order = np.argsort(h[..., 0].ravel(), kind="stable")
        red = (l & 0xFF).ravel()[order].astype(np.int64)
        assert (np.diff(red) >= 0).all()
    with pytest.raises(ptcore.PtError) as e:
        r.RenderBatch([(17, items[0][1], cornell.view)])
    assert e.value.status == 1
    with pytest.raises(ptcore.PtError) as e:
        r.RenderBatch([])
    assert e.value.status == 1


def test_add_remove_scene_slots(two_scenes, cornell, spheres):
    r, sid = two_scenes
    extra = r.AddScene(cornell)
    assert extra not in (0, sid) and 1 <= extra < 64
    st = dict(width=16, height=16, samples_per_pixel=4, path_depth=4, use_nee=1, seed=1)
    a, _ = r.RenderBatch([(extra, st, cornell.view)])
    b, _ = r.RenderBatch([(0, st, cornell.view)])
    assert np.allclose(a[0], b[0], rtol=1e-4, atol=1e-5)
    r.RemoveScene(extra)
    with pytest.raises(ptcore.PtError) as e:
        r.RenderBatch([(extra, st, cornell.view)])
    assert e.value.status == 1
    with pytest.raises(ptcore.PtError) as e:
        r.RemoveScene(extra)
    assert e.value.status == 1
    again = r.AddScene(cornell)
    assert again == extra
    r.RemoveScene(again)


def test_batch_outputs_are_optional(two_scenes, cornell, spheres):
    """The HDR and RGBA8 outputs of pt_render_batch are each optional: a view
    rendered for its RGBA8 image alone gets the same ray counts, and an image
    that tonemaps the HDR image rendered with it."""
    r, sid = two_scenes
    it = _items(cornell, spheres, sid)[0]
    (hdr,), st, (rgba,) = r.RenderBatch([it], rgba=True)
    (none,), st2, (rgba2,) = r.RenderBatch([it], rgba=True, hdr=False)
    assert none is None
    assert (st["extension_rays"], st["shadow_rays"]) == (st2["extension_rays"], st2["shadow_rays"])
    # float-atomic order differs between runs: compare channels to one 8-bit step
    a = rgba.view(np.uint8).reshape(-1, 4).astype(int)
    b = rgba2.view(np.uint8).reshape(-1, 4).astype(int)
    assert np.abs(a - b).max() <= 1
    assert (a[:, 3] == 255).all()
    assert hdr.shape == (32, 32, 4) and np.isfinite(hdr).all()
