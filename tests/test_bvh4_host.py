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


def _decode(q):
    """(N, 16) uint32 quantized nodes -> lo, hi (N, 4, 3) float32 and child refs,
    in float32 arithmetic: origin + float(byte) * 2^(e - 127)."""
    origin = q[:, 0:3].view(np.float32)
    e = np.stack([(q[:, 3] >> (8 * a)) & 255 for a in range(3)], 1)
    scale = (e.astype(np.uint32) << 23).view(np.float32)
    byte = lambda rows: np.stack([(rows >> (8 * i)) & 255 for i in range(4)], 1).astype(np.float32)   # (N, 4, 3)
    lo = origin[:, None, :] + byte(q[:, 4:7]) * scale[:, None, :]
    hi = origin[:, None, :] + byte(q[:, 7:10]) * scale[:, None, :]
    return lo.astype(np.float32), hi.astype(np.float32), scale, q[:, 10:14]


@pytest.mark.parametrize("make", [
    lambda: scenes.spheres_scene(count=6, subdivisions=4, lights=64, seed=3),
    scenes.cornell_box,
    # tiny triangles far from the origin: scale steps below the coordinates' ulp
    lambda: scenes.Scene(name="far", positions=(np.random.default_rng(4).uniform(-1e-3, 1e-3, (500, 3, 3)) +
                                                np.float32([1e4, -3e4, 7e3])).astype(np.float32),
                         normals=np.zeros((500, 3, 3), np.float32), tangents=np.zeros((500, 3, 4), np.float32),
                         material_ids=np.zeros(500, np.uint32), materials=np.stack([scenes.material(1)])),
])
def test_quantized_nodes_contain_the_fp32_boxes(make):
    """Every decoded child box contains the fp32 child box; each byte is the
    tightest one (one step further would cut into the box) and each scale the
    smallest power of two whose byte 255 reaches the node's upper corner."""
    s = make()
    nodes, _, _ = ptcore.build_bvh4(s.positions)
    q = ptcore.quantize_bvh4(nodes)
    lo, hi, scale, child = _decode(q)
    assert np.array_equal(child, nodes[:, 24:28])
    f = nodes[:, :24].view(np.float32)
    flo = np.stack([f[:, 0:4], f[:, 8:12], f[:, 16:20]], 2)        # (N, 4, 3)
    fhi = np.stack([f[:, 4:8], f[:, 12:16], f[:, 20:24]], 2)
    valid = (nodes[:, 24:28] != 0xFFFFFFFF)[:, :, None] & np.ones((1, 1, 3), bool)
    assert (lo[valid] <= flo[valid]).all() and (hi[valid] >= fhi[valid]).all()
    # tightness: the next byte up (down) decodes past the fp32 bound
    origin = q[:, 0:3].view(np.float32)[:, None, :]
    bl = np.stack([(q[:, 4:7] >> (8 * i)) & 255 for i in range(4)], 1)
    bh = np.stack([(q[:, 7:10] >> (8 * i)) & 255 for i in range(4)], 1)
    s3 = scale[:, None, :]
    up = origin + np.minimum(bl + 1, 255).astype(np.float32) * s3
    down = origin + np.maximum(bh.astype(np.int64) - 1, 0).astype(np.float32) * s3
    assert ((up > flo) | (bl == 255))[valid].all()
    assert ((down < fhi) | (bh == 0))[valid].all()
    # scale: byte 255 reaches the node's corner, half the scale would not
    node_hi = np.where(valid, fhi, -np.inf).max(1)
    o2 = q[:, 0:3].view(np.float32)
    assert (o2 + np.float32(255) * scale >= node_hi).all()
    e = (scale.view(np.uint32) >> 23)
    half = ((e - 1).astype(np.uint32) << 23).view(np.float32)
    assert ((o2 + np.float32(255) * half < node_hi) | (e == 1)).all()
    # the byte grid is fine: decoded boxes exceed the fp32 ones by at most one step (plus rounding)
    slack = np.maximum(s3, np.spacing(np.abs(origin)) * 2)
    assert ((flo - lo)[valid] <= (slack + 0 * lo)[valid] * 1.0001).all()


@pytest.mark.parametrize("make", [scenes.cornell_box, lambda: scenes.spheres_scene(count=4, subdivisions=4, lights=64)])
def test_quantized_any_hit_walk(make, oracle_lib):
    """The oracle's walk of the quantized BVH4 in the any-hit kernel's order
    (whose per-ray visits the roofline's bytes count): occlusion equal to brute
    force on random segments and on shadow segments to the emitters; a ray
    missing the root box costs one node and no triangle; an unoccluded ray,
    which cannot stop early, tests every triangle of the leaves it reaches at
    least as often as the closest-hit walk of the fp32 tree, whose boxes it
    contains."""
    s = make()
    nodes, order, depth = ptcore.build_bvh4(s.positions)
    q = ptcore.quantize_bvh4(nodes)
    o = oracle_lib.OracleScene(s)
    lo, hi = s.bounds()
    rays = scenes.random_rays(20000, lo - 1, hi + 1, 78)
    rays[:, 7] = np.random.default_rng(3).uniform(0.01, float(np.linalg.norm(hi - lo)), len(rays)).astype(np.float32)
    occ, nv, nt = o.bvh4q_any(q, order, rays)
    assert np.array_equal(occ, o.any(rays, brute=True))
    assert 0.05 < occ.mean() < 0.95 and nv.min() >= 1
    sh = scenes.shadow_rays(s, 20000, 79)
    occ_s, _, _ = o.bvh4q_any(q, order, sh)
    assert np.array_equal(occ_s, o.any(sh, brute=True))
    away = np.zeros((1, 8), np.float32)
    away[0, :3] = hi + 5
    away[0, 3], away[0, 4:7], away[0, 7] = 1e-4, 0.57735026, 1e34
    occ1, nv1, nt1 = o.bvh4q_any(q, order, away)
    assert occ1[0] == 0 and nv1[0] == 1 and nt1[0] == 0
    free = occ == 0
    _, cnv, cnt, _ = o.bvh4_closest(nodes, order, rays[free])
    assert (nt[free] >= cnt).all() and (nv[free] >= cnv).all()


def _decode8(q):
    """(N, 32) uint32 rows of 8-wide compressed nodes -> decoded child boxes lo,
    hi (N, 8, 3) and per node and slot the child: ("inner", node index),
    ("leaf", first position, count) or None, by the layout's own rules (inner
    children consecutive from the first inner child in slot order, leaf
    triangles from the first triangle, slot s ending at byte s of the ends)."""
    origin = q[:, 0:3].copy().view(np.float32)
    e = np.stack([(q[:, 3] >> (8 * a)) & 255 for a in range(3)], 1).astype(np.uint32)
    scale = (e << 23).view(np.float32)
    lo = np.zeros((len(q), 8, 3), np.float32)
    hi = np.zeros((len(q), 8, 3), np.float32)
    empty = np.ones((len(q), 8), bool)
    for a in range(3):
        for c in range(8):
            bl = (q[:, 4 + 2 * a + c // 4] >> (8 * (c % 4))) & 255
            bh = (q[:, 10 + 2 * a + c // 4] >> (8 * (c % 4))) & 255
            lo[:, c, a] = origin[:, a] + bl.astype(np.float32) * scale[:, a]
            hi[:, c, a] = origin[:, a] + bh.astype(np.float32) * scale[:, a]
            empty[:, c] &= (bl == 255) & (bh == 0)
    kids = []
    for i in range(len(q)):
        imask, base, tri_base = int(q[i, 3]) >> 24, int(q[i, 16]), int(q[i, 17])
        ends = [(int(q[i, 18 + k // 4]) >> (8 * (k % 4))) & 255 for k in range(8)]
        row, rank = [], 0
        for k in range(8):
            start = ends[k - 1] if k else 0
            if imask >> k & 1:
                assert not empty[i, k] and ends[k] == start
                row.append(("inner", base + rank))
                rank += 1
            elif ends[k] > start:
                assert not empty[i, k]
                row.append(("leaf", tri_base + start, ends[k] - start))
            else:
                assert empty[i, k]
                row.append(None)
        kids.append(row)
    return lo, hi, kids


def _leaf_sets4(nodes, order):
    sets = set()
    for i in range(len(nodes)):
        for k in range(4):
            c = int(nodes[i, 24 + k])
            if c != 0xFFFFFFFF and c & 0x80000000:
                first, cnt = (c & 0x7FFFFFFF) >> 4, (c & 15) + 1
                sets.add(frozenset(order[first:first + cnt].tolist()))
    return sets


@pytest.mark.parametrize("make", [scenes.cornell_box, lambda: scenes.spheres_scene(count=4, subdivisions=4, lights=64)])
def test_bvh8_quantized_structure_and_walk(make, oracle_lib):
    """The 8-wide compressed tree: breadth-first numbering with each node's
    inner children consecutive and after it, every node reached once, depth as
    reported; the leaves of the BVH4 build regrouped (equal triangle sets), every
    triangle in exactly one leaf whose decoded box holds it; octant slots; the
    oracle's walk in the any-hit kernel's order agrees with brute force on
    random and shadow segments, keeps at most one stack entry per level and
    visits fewer nodes than the 4-wide quantized walk."""
    s = make()
    q8, order8, depth8 = ptcore.build_bvh8_quantized(s.positions)
    nodes, order, depth = ptcore.build_bvh4(s.positions)
    T = len(s.positions)
    assert sorted(order8.tolist()) == list(range(T)) and 1 <= depth8 <= depth
    assert (q8[:, 20:32] == 0).all()
    lo, hi, kids = _decode8(q8)
    P = s.positions[order8].reshape(T, 3, 3)
    seen = np.zeros(T, np.int32)
    reached = np.zeros(len(q8), np.int32)
    reached[0] = 1
    level = np.zeros(len(q8), np.int32)
    level[0] = 1
    leaf_sets = set()
    for i in range(len(q8)):      # breadth first: parents before children
        assert reached[i] == 1
        for k, kid in enumerate(kids[i]):
            if kid is None:
                continue
            if kid[0] == "inner":
                c = kid[1]
                assert i < c < len(q8)
                reached[c] += 1
                level[c] = level[i] + 1
            else:
                _, first, cnt = kid
                seen[first:first + cnt] += 1
                pts = P[first:first + cnt].reshape(-1, 3)
                assert (pts >= lo[i, k]).all() and (pts <= hi[i, k]).all()
                leaf_sets.add(frozenset(order8[first:first + cnt].tolist()))
    assert (seen == 1).all() and (reached == 1).all() and level.max() == depth8
    assert leaf_sets == _leaf_sets4(nodes, order)
    # octant slots: slot o holds a child far upstream along octant o's diagonal
    # (o's bit a set = negative component a).  Over the nodes with several
    # children, the child of slot o is mostly the one with the least
    # projection, and never on average above the mean: a flipped octant
    # convention or an unsorted assignment fails one of these.
    first_ok, below_mean = [], []
    for i in range(len(q8)):
        occupied = [k for k in range(8) if kids[i][k] is not None]
        if len(occupied) < 3:
            continue
        cen = 0.5 * (lo[i] + hi[i])
        for oc in occupied:
            d = np.array([-1.0 if oc >> a & 1 else 1.0 for a in range(3)])
            proj = {k: float(cen[k] @ d) for k in occupied}
            first_ok.append(proj[oc] <= min(proj.values()) + 1e-6)
            below_mean.append(proj[oc] - np.mean(list(proj.values())))
    assert len(first_ok) > 10 and np.mean(first_ok) > 0.5 and np.mean(below_mean) < 0
    o = oracle_lib.OracleScene(s)
    blo, bhi = s.bounds()
    rays = scenes.random_rays(20000, blo - 1, bhi + 1, 88)
    rays[:, 7] = np.random.default_rng(4).uniform(0.01, float(np.linalg.norm(bhi - blo)), len(rays)).astype(np.float32)
    sh = scenes.shadow_rays(s, 20000, 89)
    q4 = ptcore.quantize_bvh4(nodes)
    for r in (rays, sh):
        occ, nv, nt, ms = o.bvh8q_any(q8, order8, r)
        assert np.array_equal(occ, o.any(r, brute=True))
        assert ms.max() <= depth8 and nv.min() >= 1
        _, nv4, _ = o.bvh4q_any(q4, order, r)
        assert nv.mean() < nv4.mean()
    # closest hit over the same nodes (pt_tuning.bvh_quantized bit 3): the (t, prim)
    # minimum of brute force bit for bit, whatever the visit order
    full = rays.copy()
    full[:, 7] = 1e34
    for r in (rays, full):
        hits, nv, _, ms = o.bvh8q_closest(q8, order8, r)
        assert np.array_equal(hits, o.closest(r, brute=True))
        assert ms.max() <= depth8 and nv.min() >= 1


def test_bvh8_quantized_tiny_scenes():
    """One triangle: a root whose slot 0 is a leaf of it; no triangles: a root
    with no child."""
    one = np.array([[[0, 0, 0], [1, 0, 0], [0, 1, 0]]], np.float32)
    q, order, depth = ptcore.build_bvh8_quantized(one)
    assert len(q) == 1 and depth == 1 and order.tolist() == [0]
    _, _, kids = _decode8(q)
    assert kids[0][0] == ("leaf", 0, 1) and all(k is None for k in kids[0][1:])
    q, order, depth = ptcore.build_bvh8_quantized(np.zeros((0, 3, 3), np.float32))
    assert len(q) == 1 and len(order) == 0
    _, _, kids = _decode8(q)
    assert all(k is None for k in kids[0])

def _clusters(n_clusters=40, copies=16, seed=5):
    """Clusters of near-identical triangles, which the SAH keeps as leaves of
    `copies` triangles (a split cannot shrink their boxes)."""
    rng = np.random.default_rng(seed)
    base = rng.uniform(0, 10, (n_clusters, 1, 3, 3)).astype(np.float32)
    P = np.repeat(base, copies, axis=1).reshape(-1, 3, 3)
    P = (P + rng.uniform(-1e-3, 1e-3, P.shape)).astype(np.float32)
    T = len(P)
    tan = np.zeros((T, 3, 4), np.float32)
    tan[..., 0] = tan[..., 3] = 1.0
    nrm = np.zeros((T, 3, 3), np.float32)
    nrm[..., 2] = 1.0
    return scenes.Scene(name="clusters", positions=P, normals=nrm, tangents=tan,
                        material_ids=np.zeros(T, np.uint32), materials=np.stack([scenes.material(1)]))


def test_bvh8_sixteen_triangle_leaves(oracle_lib):
    """Leaves of the largest size (16 triangles, so a node's triangle ends reach
    128): the decoded ranges cover every triangle once and the walk equals brute
    force."""
    s = _clusters()
    q8, order8, depth8 = ptcore.build_bvh8_quantized(s.positions, 16)
    _, _, kids = _decode8(q8)
    counts = [k[2] for row in kids for k in row if k and k[0] == "leaf"]
    assert max(counts) == 16 and sum(counts) == s.triangle_count
    ends = [(int(q8[i, 18 + k // 4]) >> (8 * (k % 4))) & 255 for i in range(len(q8)) for k in range(8)]
    assert max(ends) > 64
    o = oracle_lib.OracleScene(s)
    lo, hi = s.bounds()
    rays = scenes.random_rays(20000, lo - 1, hi + 1, 90)
    rays[:, 7] = np.random.default_rng(6).uniform(0.01, 20.0, len(rays)).astype(np.float32)
    occ, _, nt, _ = o.bvh8q_any(q8, order8, rays)
    assert np.array_equal(occ, o.any(rays, brute=True)) and 0.01 < occ.mean() < 0.99 and nt.max() >= 16
