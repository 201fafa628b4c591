"""The tuning space of the traversal: node width 2 (64-byte nodes) or 4
(128-byte nodes, 64-byte quantized ones, or for any hit 128-byte 8-wide
quantized ones) and block sizes 64..512.  Every
point must return the same closest hits, bit for bit (the (t, prim) minimum
does not depend on the tree or on how conservative its boxes are), and the
same images."""
import numpy as np
import pytest

import ptcore
import scenes

pytestmark = pytest.mark.gpu

# (node width, block, bvh_quantized: bit 0 closest hit, bit 1 any hit on 64-byte quantized nodes,
#  bit 2 any hit on 128-byte 8-wide quantized nodes, bit 3 closest hit on them as well)
POINTS = [(4, 256, 0), (4, 512, 0), (2, 64, 2), (2, 256, 0), (2, 512, 2), (4, 64, 0), (4, 128, 2),
          (4, 256, 3), (4, 64, 1), (4, 512, 3), (4, 256, 2), (4, 256, 6), (4, 512, 7), (4, 64, 4),
          (2, 256, 6), (4, 256, 14), (4, 128, 12), (4, 64, 15)]


@pytest.fixture(scope="module")
def rays():
    return scenes.random_rays(1 << 19, (-1, 0, -1), (1, 2, 1), 301)


@pytest.fixture(scope="module")
def brute(cornell_oracle, rays):
    return cornell_oracle.closest(rays, brute=True)


@pytest.mark.parametrize("width,block,quantized", POINTS)
def test_tuning_point_is_bit_exact(cornell, cornell_oracle, rays, brute, width, block, quantized):
    with ptcore.Renderer() as r:
        r.set_tuning(bvh_width=width, trace_block=block, bvh_quantized=quantized)
        r.InitializeRenderingBuffers(32, 32)
        r.InitializeSceneBuffers(cornell)
        info = r.bvh_info()
        assert info["node_width"] == width
        assert (info["bvh8_node_count"] > 0) == bool(quantized & 4 and width == 4)
        assert np.array_equal(r.trace_closest(rays), brute)
        sh = scenes.shadow_rays(cornell, 65536, 303)
        assert np.array_equal(r.trace_any(sh), cornell_oracle.any(sh))
        q, counts = r.classify(rays[:65536])
        assert np.array_equal(q, cornell_oracle.classify(rays[:65536]))


def test_width2_tree_is_a_binary_bvh(cornell):
    with ptcore.Renderer() as r:
        r.set_tuning(bvh_width=2)
        r.InitializeRenderingBuffers(32, 32)
        r.InitializeSceneBuffers(cornell)
        nodes, order = r.bvh4()
        info = r.bvh_info()
    child = nodes[:, 24:28]
    assert (child[:, 2:] == ptcore.INVALID).all()          # slots 2 and 3 never used
    assert sorted(order.tolist()) == list(range(cornell.triangle_count))
    assert info["stack_total"] >= info["max_depth"] + 2


def test_images_equal_across_tuning_points(cornell):
    st = dict(width=32, height=32, samples_per_pixel=16, path_depth=8, use_nee=1, seed=5)
    imgs, rays = [], []
    for width, block, quantized in (POINTS[0], POINTS[2], POINTS[7]):
        with ptcore.Renderer() as r:
            r.set_tuning(bvh_width=width, trace_block=block, bvh_quantized=quantized)
            r.InitializeRenderingBuffers(32, 32)
            r.InitializeSceneBuffers(cornell)
            s = r.RenderFrame(st, cornell.view)
            imgs.append(r.GetFrameHdr(32, 32))
            rays.append((s["extension_rays"], s["shadow_rays"]))
    assert len(set(rays)) == 1
    for im in imgs[1:]:
        assert np.allclose(im, imgs[0], rtol=1e-4, atol=1e-5)


def test_images_equal_across_traversal_grids(cornell):
    """The persistent grids of the two traversals (the automatic one, whose
    any-hit grid leaves a wave per SIMD to the concurrent closest-hit launch,
    full grids, and a single block per CU each) render the same frame."""
    st = dict(width=32, height=32, samples_per_pixel=16, path_depth=8, use_nee=1, seed=6)
    imgs, rays = [], []
    with ptcore.Renderer() as r:
        r.InitializeRenderingBuffers(32, 32)
        r.InitializeSceneBuffers(cornell)
        for closest, any_hit in ((0, 0), (6, 6), (6, 1), (1, 1), (0, 3)):
            r.set_tuning(trace_blocks_per_cu=closest, shadow_blocks_per_cu=any_hit)
            s = r.RenderFrame(st, cornell.view)
            imgs.append(r.GetFrameHdr(32, 32))
            rays.append((s["extension_rays"], s["shadow_rays"]))
    assert len(set(rays)) == 1
    for im in imgs[1:]:
        assert np.allclose(im, imgs[0], rtol=1e-4, atol=1e-5)


def test_width2_production_scene_hits(oracle_lib):
    """983,042 triangles with 64-byte nodes, 128-byte nodes and quantized
    64-byte nodes: equal hits and occlusion, and equal to the oracle's on a
    seeded subset (ties between coplanar triangles aside)."""
    s = scenes.production_scene()
    rays = scenes.random_rays(1 << 20, (-10, -10, -10), (110, 110, 110), 302)
    with ptcore.Renderer() as r:
        r.set_tuning(bvh_width=2)
        r.InitializeRenderingBuffers(32, 32)
        r.InitializeSceneBuffers(s)
        hits = r.trace_closest(rays)
        r.set_tuning(bvh_width=4, bvh_quantized=0)
        r.InitializeSceneBuffers(s)
        hits4 = r.trace_closest(rays)
        sh = scenes.shadow_rays(s, 1 << 18, 304)
        occ4 = r.trace_any(sh)
        r.set_tuning(bvh_quantized=3)
        r.InitializeSceneBuffers(s)
        hits4q = r.trace_closest(rays)
        occ4q = r.trace_any(sh)
    assert np.array_equal(hits, hits4) and np.array_equal(hits4, hits4q)
    assert np.array_equal(occ4, occ4q)
    o = oracle_lib.OracleScene(s)
    sub = np.random.default_rng(1).choice(len(rays), 16384, replace=False)
    ref = o.closest(rays[sub])
    diff = (hits[sub] != ref).any(1)
    assert diff.mean() < 1e-3
    assert np.array_equal(hits[sub][diff, 1], ref[diff, 1])


def test_invalid_tuning_is_rejected(gpu_renderer):
    for bad in (dict(bvh_width=3), dict(trace_block=1024), dict(trace_block=96), dict(bvh_quantized=8), dict(bvh_quantized=16),
                dict(trace_blocks_per_cu=65), dict(shadow_blocks_per_cu=65)):
        with pytest.raises(ptcore.PtError) as e:
            gpu_renderer.set_tuning(**bad)
        assert e.value.status == 1
