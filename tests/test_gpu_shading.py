"""GPU shading against the oracle and the closed forms (MaterialEval.cu:25-912).

The core and the oracle draw every random number of a sample from the same
addressed stream and evaluate the same float operations (contraction off,
IEEE division and square root, the same sampling trig), so each sample's path
is the same on both: at one sample per pixel every pixel agrees to the
rounding of the float additions into it, and ray counts are equal.  The
closed-form scenes hold the GPU to the same mathematics the oracle is held to.
"""
import math

import numpy as np
import pytest

import closed_form_scenes as cf
import ptcore
import scenes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu():
    r = ptcore.Renderer()
    r.InitializeRenderingBuffers(64, 64)
    yield r
    r.close()


def _render(r, scene, st):
    r.InitializeSceneBuffers(scene)
    stats = r.RenderFrame(st, scene.view)
    return r.GetFrameHdr(st["width"], st["height"]).reshape(-1, 4), stats


SCENES = {
    "cornell": scenes.cornell_box,
    "specular_room": scenes.specular_room,
    "plates": cf.plates,
    "corridor": cf.corridor,
    "ggx_floor": lambda: cf.ggx_floor(((0.95, 0.64, 0.54), 0.3, 0.7, 1.5)),
    "sphere_furnace": cf.sphere_furnace,
}


@pytest.mark.parametrize("nee", [1, 0])
@pytest.mark.parametrize("name", list(SCENES))
def test_one_sample_per_pixel_equals_the_oracle(gpu, oracle_lib, name, nee):
    """One sample per pixel: each pixel is one sample's radiance, summed in
    float atomics on the GPU and in double in the oracle; they agree to 1e-5
    relative everywhere and the ray counts are equal."""
    s = SCENES[name]()
    st = dict(width=96, height=96, samples_per_pixel=1, path_depth=8, use_nee=nee, rr_start_depth=3, seed=17)
    hdr, stats = _render(gpu, s, st)
    ref, ext, sh = oracle_lib.OracleScene(s).render(st, s.view, tprim=True)
    assert stats["extension_rays"] == ext and stats["shadow_rays"] == sh
    bad = ~np.isclose(hdr[:, :3], ref[:, :3], rtol=1e-5, atol=1e-6).all(1)
    assert not bad.any(), (int(bad.sum()), hdr[bad][:4], ref[bad][:4])
    assert (ref[:, :3] > 0).mean() > (0.3 if nee else 0.01)


def _relmse(a, b, eps=1e-2):
    return float(np.mean((a - b) ** 2 / (b ** 2 + eps)))


def _first_hit_materials(scene, oracle_scene, w, h):
    """Material model seen through each pixel's centre."""
    o, llc, hor, ver = cf.camera_basis64(scene.view, w / h)
    x, y = np.meshgrid((np.arange(w) + 0.5) / w, (np.arange(h) + 0.5) / h)
    d = llc[None, None] + x[..., None] * hor + y[..., None] * ver - o
    d /= np.linalg.norm(d, axis=-1, keepdims=True)
    rays = np.zeros((w * h, 8), np.float32)
    rays[:, 0:3] = o
    rays[:, 3] = 1e-4
    rays[:, 4:7] = d.reshape(-1, 3)
    rays[:, 7] = 1e34
    prim = oracle_scene.closest(rays, tprim=True)[:, 0]
    model = np.full(w * h, -1)
    hit = prim != ptcore.INVALID
    model[hit] = scene.materials["model"][scene.material_ids[prim[hit]]]
    return model


@pytest.mark.parametrize("nee", [1, 0])
def test_specular_room_matches_the_oracle_per_material(gpu, oracle_lib, nee):
    """The room whose floor, walls and boxes are GGX and mirror, 64 x 64 at 256
    spp: relMSE and a 3-sigma mean test per first-hit material (GGX, mirror,
    diffuse), against the oracle under both tie rules; ray counts equal."""
    s = scenes.specular_room()
    o = oracle_lib.OracleScene(s)
    w = h = 64
    st = dict(width=w, height=h, samples_per_pixel=256, path_depth=8, use_nee=nee, rr_start_depth=3, seed=23)
    hdr, stats = _render(gpu, s, st)
    ref, ext, sh = o.render(st, s.view, tprim=True)
    assert stats["extension_rays"] == ext and stats["shadow_rays"] == sh
    model = _first_hit_materials(s, o, w, h)
    for m, min_px in ((scenes.MODEL_GGX, 1500), (scenes.MODEL_MIRROR, 600), (scenes.MODEL_DIFFUSE, 100)):
        mask = model == m
        assert mask.sum() >= min_px, (m, int(mask.sum()))
        g, r = hdr[mask, :3], ref[mask, :3]
        rm = _relmse(g, r)
        assert rm < 1e-8, (m, rm)
        se = math.sqrt(r.mean(1).var() / mask.sum())
        assert abs(g.mean() - r.mean()) < 3 * se + 1e-6
    # the reference's tie order changes no image statistic
    ref_order, _, _ = o.render(st, s.view)
    assert _relmse(hdr[:, :3], ref_order[:, :3]) < 1e-3


@pytest.mark.parametrize("nee, depth, rr", [(1, 8, 100), (0, 8, 100), (1, 1, 100), (0, 2, 1), (1, 8, 1)])
def test_gpu_diffuse_plates_match_form_factor(gpu, nee, depth, rr):
    s = cf.plates(0.6, 3.0)
    st = dict(width=32, height=32, samples_per_pixel=256, path_depth=depth, use_nee=nee, rr_start_depth=rr, seed=7)
    hdr, stats = _render(gpu, s, st)
    cf.check_plates(hdr, s.view, 32, 32, 0.6, 3.0, nee, depth)


@pytest.mark.parametrize("depth, spp, rr", [(8, 16, 100), (4, 16, 100), (16, 256, 1)])
def test_gpu_mirror_corridor_power_series(gpu, depth, spp, rr):
    s = cf.corridor()
    st = dict(width=64, height=16, samples_per_pixel=spp, path_depth=depth, use_nee=1, rr_start_depth=rr, seed=3)
    hdr, stats = _render(gpu, s, st)
    assert stats["shadow_rays"] == 0
    cf.check_corridor(hdr, s.view, 64, 16, depth, spp, rr)


@pytest.mark.parametrize("nee, depth, rr", [(1, 1, 100), (1, 8, 100), (0, 3, 100), (0, 8, 100), (1, 8, 1), (0, 8, 2)])
def test_gpu_diffuse_sphere_furnace_is_a_geometric_series(gpu, nee, depth, rr):
    """Inside a sphere a quarter emissive, diffuse pixels read rho f Le
    (1 + q + ... + q^(n-1)), q = rho (1 - f), n = depth with NEE and depth - 1
    without; emissive pixels read Le."""
    rho, le = 0.8, 2.0
    s = cf.sphere_furnace(rho, le)
    w = h = 32
    st = dict(width=w, height=h, samples_per_pixel=256, path_depth=depth, use_nee=nee, rr_start_depth=rr, seed=4)
    hdr, _ = _render(gpu, s, st)
    mask = cf.furnace_pixels(s, gpu.trace_closest, w, h)
    assert mask.sum() > 600
    cf.check_furnace(hdr, mask, rho, le, cf.furnace_fraction(s), depth, nee)
    emit = cf.furnace_pixels(s, gpu.trace_closest, w, h, material=1)
    assert emit.sum() > 50 and np.allclose(hdr[emit, :3], le, rtol=1e-5)


@pytest.mark.parametrize("material", [((1.0, 1.0, 1.0), 0.5, 1.0, 1.5), ((0.8, 0.4, 0.2), 0.3, 0.5, 1.5),
                                      ((0.7, 0.7, 0.7), 0.9, 0.0, 1.5)])
@pytest.mark.parametrize("nee", [1, 0])
def test_gpu_ggx_floor_matches_quadrature(gpu, material, nee):
    s = cf.ggx_floor(material)
    st = dict(width=16, height=16, samples_per_pixel=1024, path_depth=2, use_nee=nee, rr_start_depth=100, seed=11)
    hdr, _ = _render(gpu, s, st)
    cf.check_ggx_floor(hdr, s.view, 16, 16, material)


def test_gpu_tonemap_equals_the_oracle(gpu, oracle_lib):
    """PostprocessKernel (Image.cu:34-54): the GPU's RGBA8 frame equals the
    oracle's tonemap of the GPU's own radiance / spp within one step of 255
    (powf's last ulp), and exactly for nearly all channels."""
    s = scenes.specular_room()
    st = dict(width=64, height=64, samples_per_pixel=16, path_depth=8, use_nee=1, rr_start_depth=3, seed=3)
    hdr, _ = _render(gpu, s, st)
    rgba = gpu.GetFrame(64, 64).reshape(-1)
    want = oracle_lib.tonemap(hdr)
    a = np.stack([(rgba >> (8 * k)) & 0xFF for k in range(4)], 1).astype(int)
    b = np.stack([(want >> (8 * k)) & 0xFF for k in range(4)], 1).astype(int)
    assert (a[:, 3] == 255).all() and np.abs(a - b).max() <= 1
    assert (a == b).mean() > 0.999
    assert len(np.unique(a[:, :3])) > 64
