"""CPU checks of the core's BVH4 build (host code of libptcore.so, no GPU):
structure, containment, and closest hits of a walk in the GPU's visit order
against brute force."""
import numpy as np
import pytest

import ptcore
import scenes


def _check_structure(nodes, order, depth, positions):
    T = len(positions)
    assert sorted(order.tolist()) == list(range(T))
    f = nodes[:, :24].view(np.float32)
    child = nodes[:, 24:28]
    seen = np.zeros(T, np.int32)
    P = positions[order]
    stack = [(0, 1)]
    max_level = 0
    while stack:
        i, level = stack.pop()
        max_level = max(max_level, level)
        for k in range(4):
            c = child[i, k]
            if c == 0xFFFFFFFF:
                continue
            lo = np.array([f[i, 0 + k], f[i, 8 + k], f[i, 16 + k]])
            hi = np.array([f[i, 4 + k], f[i, 12 + k], f[i, 20 + k]])
            if c & 0x80000000:
                first, cnt = (c & 0x7FFFFFFF) >> 4, (c & 15) + 1
                seen[first:first + cnt] += 1
                pts = P[first:first + cnt].reshape(-1, 3)
                assert (pts > lo).all() and (pts < hi).all()   # padded: strictly inside
            else:
                assert 0 < c < len(nodes)
                stack.append((int(c), level + 1))
    assert (seen == 1).all()
    assert max_level == depth


@pytest.mark.parametrize("make", [scenes.cornell_box, lambda: scenes.spheres_scene(count=4, subdivisions=4, lights=64)])
def test_bvh4_structure_and_walk(make, oracle_lib):
    s = make()
    nodes, order, depth = ptcore.build_bvh4(s.positions)
    _check_structure(nodes, order, depth, s.positions)
    o = oracle_lib.OracleScene(s)
    lo, hi = s.bounds()
    rays = scenes.random_rays(20000, lo - 1, hi + 1, 77)
    hits, nv, nt, ms = o.bvh4_closest(nodes, order, rays)
    assert np.array_equal(hits, o.closest(rays, brute=True))
    assert ms.max() <= 3 * depth + 2      # the kernel's stack bound
    assert nv.min() >= 1


def test_bvh4_empty_and_single_triangle():
    nodes, order, depth = ptcore.build_bvh4(np.zeros((0, 3, 3), np.float32))
    assert len(nodes) == 1 and len(order) == 0
    assert (nodes[0, 24:28] == 0xFFFFFFFF).all()
    tri = np.array([[[0, 0, 0], [1, 0, 0], [0, 1, 0]]], np.float32)
    nodes, order, depth = ptcore.build_bvh4(tri)
    assert order.tolist() == [0] and nodes[0, 24] & 0x80000000


def test_bvh4_rejects_non_finite():
    tri = np.array([[[0, 0, 0], [np.nan, 0, 0], [0, 1, 0]]], np.float32)
    with pytest.raises(ptcore.PtError):
        ptcore.build_bvh4(tri)
