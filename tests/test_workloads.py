"""CPU checks of the synthetic workloads: the scaled config-1 room, the
batching workload and the serving requests."""
import math

import numpy as np

import scenes


def _areas(s):
    P = s.positions.astype(np.float64)
    return 0.5 * np.linalg.norm(np.cross(P[:, 1] - P[:, 0], P[:, 2] - P[:, 0]), axis=1)


def _geo_normals(s):
    P = s.positions.astype(np.float64)
    n = np.cross(P[:, 1] - P[:, 0], P[:, 2] - P[:, 0])
    return n / np.linalg.norm(n, axis=1, keepdims=True)


def test_tessellated_room_covers_the_same_surfaces():
    c1 = scenes.cornell_box()
    for k in (2, 5):
        ck = scenes.cornell_box(k)
        assert ck.triangle_count == 36 * k * k
        # per material, the same area, the same facing, the same bounds
        for m in range(len(c1.materials)):
            a1 = _areas(c1)[c1.material_ids == m].sum()
            ak = _areas(ck)[ck.material_ids == m].sum()
            assert math.isclose(a1, ak, rel_tol=1e-6), (k, m)
        assert np.allclose(ck.positions.min((0, 1)), c1.positions.min((0, 1)))
        assert np.allclose(ck.positions.max((0, 1)), c1.positions.max((0, 1)))
        # winding follows the shading normal, as in the 36-triangle room
        assert (np.einsum("ij,ij->i", _geo_normals(ck), ck.normals[:, 0].astype(np.float64)) > 0.999).all()
        assert (_areas(ck) > 0).all()


def test_cornell_scaled_hits_the_requested_size():
    for target in (10_000, 250_000):
        s = scenes.cornell_scaled(target)
        k = math.sqrt(target / 36)          # an integer k changes the count by about 2/k
        assert abs(s.triangle_count - target) / target < 1.2 / k


def test_batch_workload_shape():
    sc, views = scenes.batch_workload(count=6, seed=5, tri_range=(2_000, 20_000))
    assert len(sc) == 6
    for s in sc:
        assert 1_500 < s.triangle_count < 30_000
        assert np.isfinite(s.positions).all() and (s.material_ids < len(s.materials)).all()
        assert (s.materials["model"][s.material_ids] == scenes.MODEL_EMISSIVE).sum() == 2
    per_scene = np.bincount([i for i, _, _ in views], minlength=6)
    assert per_scene.min() >= 1 and per_scene.max() <= 4
    for _, st, v in views:
        assert (st["width"], st["height"]) in scenes.BATCH_RESOLUTIONS
        assert st["samples_per_pixel"] in scenes.BATCH_SPP
        q = np.asarray(v.rotation)
        assert math.isclose(float(np.dot(q, q)), 1.0, rel_tol=1e-6)


def test_serve_requests_are_seeded_and_in_range():
    a = [scenes.serve_request(np.random.default_rng(3), 8) for _ in range(20)]
    b = [scenes.serve_request(np.random.default_rng(3), 8) for _ in range(20)]
    assert [(x[0], x[1]) for x in a] == [(y[0], y[1]) for y in b]
    rng = np.random.default_rng(4)
    for _ in range(200):
        sid, st, v = scenes.serve_request(rng, 8)
        assert 0 <= sid < 8
        assert (st["width"], st["height"]) in scenes.SERVE_RESOLUTIONS
        assert st["samples_per_pixel"] in scenes.SERVE_SPP
        # square spp counts: the renderer keeps them as given
        assert int(math.isqrt(st["samples_per_pixel"])) ** 2 == st["samples_per_pixel"]
