"""Degenerate geometry: shared edges and vertices, duplicated and coplanar
overlapping triangles (exact ties in t, resolved to the lower primitive id),
zero-area triangles (never hit) and slivers.  Closest and any hit on the GPU
must equal the oracle's brute force bit for bit, rays aimed exactly at
vertices and edges included."""
import math

import numpy as np
import pytest

import ptcore
import scenes

pytestmark = pytest.mark.gpu


# triangle id ranges of the scene below
GRID = range(0, 128)
DUPLICATES = range(128, 168)         # copies of GRID[0:40]
ZERO_AREA = range(208, 211)


def _torture_scene():
    P, N = [], []

    def add(tris, normal=(0.0, 0.0, 1.0)):
        t = np.asarray(tris, np.float32).reshape(-1, 3, 3)
        P.append(t)
        N.append(np.broadcast_to(np.asarray(normal, np.float32), t.shape).copy())

    grid, _, _ = scenes._grid_quad((0, 0, 0), (8, 0, 0), (8, 8, 0), (0, 8, 0), (0, 0, 1), 8)
    add(grid)                                   # 128 triangles sharing edges and vertices on z = 0
    add(grid[:40])                              # exact duplicates: ties in t, u and v
    shifted = grid[40:80].copy()
    shifted[..., 0] += 0.37                     # coplanar, overlapping, other (u, v)
    shifted[..., 1] += 0.21
    add(shifted)
    add([[(1, 1, 0.5), (1, 1, 0.5), (3, 2, 0.5)],          # two equal vertices
         [(2, 2, 0.7), (4, 4, 0.7), (6, 6, 0.7)],          # collinear
         [(5, 5, 0.9), (5, 5, 0.9), (5, 5, 0.9)]])         # a point
    add([[(0, 4, 1.0), (8, 4.001, 1.0), (8, 4.0, 1.0)],    # slivers
         [(4, 0, 1.5), (4.0005, 8, 1.5), (4, 8, 1.5)]])
    tilt, _, _ = scenes._grid_quad((0, 0, 2), (8, 0, 4), (8, 8, 4), (0, 8, 2), (-0.24, 0, 0.97), 4)
    add(tilt, (-0.24, 0.0, 0.97))              # a tilted layer above
    P = np.concatenate(P)
    N = np.concatenate(N)
    T = np.zeros((len(P), 3, 4), np.float32)
    T[..., 0] = 1.0
    T[..., 3] = 1.0
    mats = np.stack([scenes.material(scenes.MODEL_DIFFUSE, color=(0.7, 0.7, 0.7)),
                     scenes.material(scenes.MODEL_EMISSIVE, emission=(5.0, 5.0, 5.0))])
    M = np.zeros(len(P), np.uint32)
    M[-1] = 1                                   # one light
    return scenes.Scene(name="degenerate", positions=P, normals=N, tangents=T, material_ids=M, materials=mats,
                        view=scenes.View(math.radians(60.0), (4.0, -6.0, 6.0),
                                         scenes.look_at_quaternion((4.0, -6.0, 6.0), (4.0, 4.0, 0.0))))


def _rays(n, seed):
    rng = np.random.default_rng(seed)
    r = scenes.random_rays(n, (-1.0, -1.0, -3.0), (9.0, 9.0, 6.0), seed)
    k = n // 4
    # straight down onto grid vertices, edge midpoints and cell diagonals
    ij = rng.integers(0, 9, (k, 2)).astype(np.float32)
    off = np.array([[0.0, 0.0], [0.5, 0.0], [0.0, 0.5], [0.5, 0.5]], np.float32)[rng.integers(0, 4, k)]
    r[:k, 0:2] = ij + off
    r[:k, 2] = 3.0
    r[:k, 4:7] = (0.0, 0.0, -1.0)
    # oblique rays through the same points
    d = np.stack([rng.uniform(-0.5, 0.5, k), rng.uniform(-0.5, 0.5, k), -np.ones(k)], 1)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    p = np.concatenate([rng.integers(0, 9, (k, 2)).astype(np.float64), np.zeros((k, 1))], 1)
    r[k:2 * k, 0:3] = p - 3.0 * d
    r[k:2 * k, 4:7] = d
    return r


@pytest.fixture(scope="module")
def torture(gpu_renderer, oracle_lib):
    s = _torture_scene()
    assert s.triangle_count == 245
    gpu_renderer.InitializeRenderingBuffers(32, 32)
    gpu_renderer.InitializeSceneBuffers(s)
    return gpu_renderer, oracle_lib.OracleScene(s), s


def test_degenerate_closest_hit_equals_brute_force(torture):
    r, o, s = torture
    rays = _rays(1 << 18, 7)
    got = r.trace_closest(rays)
    want = o.closest(rays, brute=True)
    assert np.array_equal(got, want)
    hit = got[:, 0] != ptcore.INVALID
    assert hit.mean() > 0.5
    # zero-area triangles are never hit; duplicates lose their ties to the originals
    assert not np.isin(got[hit, 0], list(ZERO_AREA)).any()
    assert not np.isin(got[hit, 0], list(DUPLICATES)).any()
    assert np.isin(got[hit, 0], list(GRID)).mean() > 0.2


def test_degenerate_any_hit_equals_brute_force(torture):
    r, o, s = torture
    rays = _rays(1 << 18, 8)
    rays[:, 7] = np.random.default_rng(3).uniform(0.1, 6.0, len(rays)).astype(np.float32)
    assert np.array_equal(r.trace_any(rays), o.any(rays, brute=True))


def test_degenerate_render_matches_oracle(torture):
    r, o, s = torture
    st = dict(width=24, height=16, samples_per_pixel=16, path_depth=4, use_nee=1, rr_start_depth=3, seed=3)
    stats = r.RenderFrame(st, s.view)
    hdr = r.GetFrameHdr(24, 16).reshape(-1, 4)
    ref, ext, sh = o.render(st, s.view)
    assert abs(stats["extension_rays"] - ext) <= 0.01 * ext and abs(stats["shadow_rays"] - sh) <= 0.01 * sh
    assert float(np.mean((hdr[:, :3] - ref[:, :3]) ** 2 / (ref[:, :3] ** 2 + 1e-2))) < 1e-3


def _dark_room():
    """The box room with its ceiling light made a white wall: no emissive
    triangle, so no light table."""
    s = scenes.cornell_box()
    mats = s.materials.copy()
    for k in range(len(mats)):
        if mats[k]["model"] == scenes.MODEL_EMISSIVE:
            mats[k] = scenes.material(scenes.MODEL_DIFFUSE, color=(0.73, 0.73, 0.73))
    s.materials = mats
    return s


@pytest.mark.parametrize("nee", [1, 0])
def test_scene_without_emitters_renders_black(oracle_lib, nee):
    """No emissive triangle: next-event estimation has nothing to sample and is
    skipped (the reference would sample an empty alias table), every pixel is
    zero, the ray counts equal the oracle's, and the alias hook refuses."""
    s = _dark_room()
    r = ptcore.Renderer()
    try:
        r.InitializeRenderingBuffers(32, 32)
        r.InitializeSceneBuffers(s)
        assert r.bvh_info()["light_count"] == 0
        st = dict(width=32, height=32, samples_per_pixel=16, path_depth=8, use_nee=nee, rr_start_depth=3, seed=4)
        stats = r.RenderFrame(st, s.view)
        hdr = r.GetFrameHdr(32, 32)
        assert np.all(hdr[..., :3] == 0.0)
        assert (r.GetFrame(32, 32) == 0xFF000000).all()
        _, ext, sh = oracle_lib.OracleScene(s).render(st, s.view)
        assert stats["extension_rays"] == ext and stats["shadow_rays"] == sh == 0
        with pytest.raises(ptcore.PtError) as e:
            r.alias_sample(np.array([0.5], np.float32))
        assert e.value.status == 2   # PT_ERROR_INVALID_STATE
    finally:
        r.close()


def test_view_that_misses_everything(oracle_lib):
    """A camera looking away from all geometry: one extension ray per sample,
    no shadow ray, a black frame."""
    s = scenes.cornell_box()
    away = scenes.View(s.view.fov_y, (0.0, 1.0, 50.0), scenes.look_at_quaternion((0.0, 1.0, 50.0), (0.0, 1.0, 100.0)))
    r = ptcore.Renderer()
    try:
        r.InitializeRenderingBuffers(24, 16)
        r.InitializeSceneBuffers(s)
        st = dict(width=24, height=16, samples_per_pixel=4, path_depth=8, use_nee=1, rr_start_depth=3, seed=9)
        stats = r.RenderFrame(st, away)
        assert stats["extension_rays"] == 24 * 16 * 4 and stats["shadow_rays"] == 0
        assert np.all(r.GetFrameHdr(24, 16)[..., :3] == 0.0)
        _, ext, sh = oracle_lib.OracleScene(s).render(st, away)
        assert (ext, sh) == (24 * 16 * 4, 0)
    finally:
        r.close()


def test_sixteen_triangle_leaves_any_hit(oracle_lib):
    """Leaves of the largest size (pt_tuning.bvh_max_leaf = 16) in the 8-wide
    any-hit tree: occlusion equal to brute force and per-ray node and triangle
    visits equal to the oracle's walk of the uploaded tree."""
    from test_bvh4_host import _clusters
    s = _clusters()
    with ptcore.Renderer() as r:
        r.set_tuning(bvh_max_leaf=16)
        r.InitializeSceneBuffers(s)
        q8, order8 = r.bvh8_quantized()
        o = oracle_lib.OracleScene(s)
        lo, hi = s.bounds()
        rays = scenes.random_rays(1 << 16, lo - 1, hi + 1, 91)
        rays[:, 7] = np.random.default_rng(7).uniform(0.01, 20.0, len(rays)).astype(np.float32)
        occ, nv, nt = r.trace_any_counts(rays)
        oocc, onv, ont, _ = o.bvh8q_any(q8, order8, rays)
        assert np.array_equal(occ, o.any(rays, brute=True))
        assert np.array_equal(occ, oocc) and np.array_equal(nv, onv) and np.array_equal(nt, ont)
        assert nt.max() >= 16
