"""The GPU BVH builder (linear BVH over Morton codes, collapsed to BVH4):
the tree it leaves on the device passes the host tree's structure checks
(every triangle in exactly one leaf, padded boxes strictly containing their
triangles, depth as reported), and traversal and rendering over it give the
same results as over the host's binned-SAH tree."""
import numpy as np
import pytest

import ptcore
import scenes
from test_bvh4_host import _check_structure

pytestmark = pytest.mark.gpu


@pytest.fixture
def gpu_builder(gpu_renderer):
    tun = gpu_renderer.get_tuning()
    gpu_renderer.set_tuning(bvh_builder=1)
    yield gpu_renderer
    gpu_renderer.set_tuning(**tun)


def _tree(r, s):
    nodes, order = r.bvh4()
    info = r.bvh_info()
    _check_structure(nodes, order, info["max_depth"], s.positions)
    assert info["max_leaf_size"] <= r.get_tuning()["bvh_max_leaf"]
    return info


@pytest.mark.parametrize("max_leaf", [1, 4, 8, 16])
def test_gpu_bvh_cornell_structure_and_hits(gpu_builder, cornell, cornell_oracle, max_leaf):
    r = gpu_builder
    r.set_tuning(bvh_max_leaf=max_leaf)
    r.InitializeRenderingBuffers(64, 64)
    r.InitializeSceneBuffers(cornell)
    _tree(r, cornell)
    rays = scenes.random_rays(1 << 18, (-1, 0, -1), (1, 2, 1), 41)
    assert np.array_equal(r.trace_closest(rays), cornell_oracle.closest(rays, brute=True))
    sh = scenes.shadow_rays(cornell, 1 << 16, 42)
    assert np.array_equal(r.trace_any(sh), cornell_oracle.any(sh, brute=True))


def test_gpu_quantized_nodes_equal_host_quantizer(gpu_builder, cornell):
    """Nodes quantized on the GPU hold the bytes the host quantizer gives for
    the same fp32 nodes: the traversal sees one format whichever side made it."""
    r = gpu_builder
    r.set_tuning(bvh_quantized=3)
    r.InitializeRenderingBuffers(32, 32)
    r.InitializeSceneBuffers(cornell)
    nodes, _ = r.bvh4()
    q_dev = r.bvh4_quantized()
    assert np.array_equal(q_dev, ptcore.quantize_bvh4(nodes))


@pytest.mark.parametrize("kind", ["one", "two", "same"])
def test_gpu_bvh_tiny_and_identical(gpu_builder, oracle_lib, kind):
    """One triangle, two, and 300 copies of one triangle (equal Morton codes)."""
    tri = np.array([[[0, 0, 0], [1, 0, 0], [0, 1, 0]]], np.float32)
    P = {"one": tri, "two": np.concatenate([tri, tri + 0.5]), "same": np.repeat(tri, 300, axis=0)}[kind]
    n = len(P)
    N = np.broadcast_to(np.array([0, 0, 1], np.float32), P.shape).copy()
    T = np.zeros((n, 3, 4), np.float32)
    T[..., 0] = 1.0
    T[..., 3] = 1.0
    s = scenes.Scene(name=kind, positions=P, normals=N, tangents=T, material_ids=np.zeros(n, np.uint32),
                     materials=np.stack([scenes.material(scenes.MODEL_DIFFUSE)]))
    r = gpu_builder
    r.InitializeSceneBuffers(s)
    _tree(r, s)
    o = oracle_lib.OracleScene(s)
    rays = scenes.random_rays(1 << 14, (-0.5, -0.5, -1.0), (1.5, 1.5, 1.0), 43)
    assert np.array_equal(r.trace_closest(rays), o.closest(rays, brute=True))


def test_gpu_bvh_production_scene_equals_host_tree(gpu_renderer):
    """983,042 triangles: hits over the GPU-built tree equal those over the
    host's tree bit for bit, and both renders of a small frame agree."""
    s = scenes.production_scene()
    r = gpu_renderer
    tun = r.get_tuning()
    rays = scenes.random_rays(1 << 20, (-10, -10, -10), (110, 110, 110), 44)
    sh = scenes.shadow_rays(s, 1 << 18, 45)
    st = dict(s.settings, width=96, height=64, samples_per_pixel=16)
    try:
        r.InitializeRenderingBuffers(96, 64)
        r.set_tuning(bvh_builder=0)
        r.InitializeSceneBuffers(s)
        host_hits, host_occ = r.trace_closest(rays), r.trace_any(sh)
        host_stats = r.RenderFrame(st, s.view)
        host_img = r.GetFrameHdr(96, 64)
        r.set_tuning(bvh_builder=1)
        r.InitializeSceneBuffers(s)
        info = _tree(r, s)
        assert info["triangle_count"] == 983042 and info["node_count"] > 983042 // 32
        assert np.array_equal(r.trace_closest(rays), host_hits)
        assert np.array_equal(r.trace_any(sh), host_occ)
        stats = r.RenderFrame(st, s.view)
        img = r.GetFrameHdr(96, 64)
        # the GPU-built tree traversed as quantized 64-byte nodes (quantized on the GPU)
        r.set_tuning(bvh_builder=1, bvh_quantized=3)
        r.InitializeSceneBuffers(s)
        assert np.array_equal(r.trace_closest(rays), host_hits)
        assert np.array_equal(r.trace_any(sh), host_occ)
    finally:
        r.set_tuning(**tun)
    assert stats["extension_rays"] == host_stats["extension_rays"]
    assert stats["shadow_rays"] == host_stats["shadow_rays"]
    assert np.allclose(img, host_img, rtol=1e-4, atol=1e-5)
