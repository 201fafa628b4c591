"""Tangent frames (pt_generate_tangents, host code of libptcore.so) against the
reference's own mikktspace (External/mikktspace/mikktspace.c compiled into
oracle/_ref, driven as TangentSpace.cpp:14-89 drives it): equal values on
smooth, seamed, mirrored, degenerate, non-manifold and random meshes; and
against the geometry on meshes whose frames are known in closed form."""
import os
import sys

import numpy as np
import pytest

import ptcore

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "..", "oracle"))
import build as oracle_build  # noqa: E402
import oracle  # noqa: E402


@pytest.fixture(scope="module")
def ref():
    if not os.path.exists(oracle.REF_MIKK) and oracle_build.build_ref() is None:
        pytest.skip("the reference's mikktspace sources are not present and oracle/_ref is not built")
    return oracle.reference_tangents


def _unit(v):
    v = np.asarray(v, np.float32)
    return (v / np.linalg.norm(v, axis=-1, keepdims=True)).astype(np.float32)


def _height_grid(n, seed, mirror=False, bumps=0.3):
    """(n+1)^2 vertices of a bumpy height field, smooth normals, planar uvs
    (u mirrored about the middle column when `mirror`), two triangles per cell."""
    rng = np.random.default_rng(seed)
    x, y = np.meshgrid(np.linspace(0, 1, n + 1), np.linspace(0, 1, n + 1), indexing="xy")
    a = rng.uniform(0.5, 3.0, 4)
    z = bumps * (np.sin(a[0] * x * 6 + a[1]) * np.cos(a[2] * y * 6 + a[3]))
    dzdx = bumps * 6 * a[0] * np.cos(a[0] * x * 6 + a[1]) * np.cos(a[2] * y * 6 + a[3])
    dzdy = -bumps * 6 * a[2] * np.sin(a[0] * x * 6 + a[1]) * np.sin(a[2] * y * 6 + a[3])
    P = np.stack([x, y, z], -1).reshape(-1, 3).astype(np.float32)
    N = _unit(np.stack([-dzdx, -dzdy, np.ones_like(x)], -1).reshape(-1, 3))
    u = np.abs(x - 0.5) if mirror else x
    U = np.stack([u, y], -1).reshape(-1, 2).astype(np.float32)
    i = np.arange(n)[:, None] * (n + 1) + np.arange(n)[None, :]
    quads = np.stack([i, i + 1, i + n + 2, i + n + 1], -1).reshape(-1, 4)
    idx = np.concatenate([quads[:, [0, 1, 2]], quads[:, [0, 2, 3]]], 1).reshape(-1)
    return P, N, U, idx.astype(np.uint32)


def _sphere(nu, nv):
    """UV sphere with a row of vertices per pole: the pole triangles are
    degenerate, the u = 0 / u = 1 seam splits vertices with equal positions."""
    th = np.linspace(0, np.pi, nv + 1)
    ph = np.linspace(0, 2 * np.pi, nu + 1)
    T, Ph = np.meshgrid(th, ph, indexing="ij")
    P = np.stack([np.sin(T) * np.cos(Ph), np.sin(T) * np.sin(Ph), np.cos(T)], -1).reshape(-1, 3).astype(np.float32)
    P[np.abs(P) < 1e-7] = 0.0
    N = P.copy()
    N[: nu + 1] = (0, 0, 1)
    N[-(nu + 1):] = (0, 0, -1)
    U = np.stack([Ph / (2 * np.pi), T / np.pi], -1).reshape(-1, 2).astype(np.float32)
    i = np.arange(nv)[:, None] * (nu + 1) + np.arange(nu)[None, :]
    quads = np.stack([i, i + nu + 1, i + nu + 2, i + 1], -1).reshape(-1, 4)
    idx = np.concatenate([quads[:, [0, 1, 2]], quads[:, [0, 2, 3]]], 1).reshape(-1)
    return P, _unit(N), U, idx.astype(np.uint32)


def _awkward():
    """Shared edges with three and four faces, a 40-face fan, duplicated
    vertices with equal attributes, zero-uv-area faces (alone and next to good
    ones), repeated and collinear positions, and a face whose uvs are mirrored
    against its neighbours'."""
    P, N, U, I = [], [], [], []

    def vert(p, n, uv):
        P.append(p)
        N.append(n)
        U.append(uv)
        return len(P) - 1

    up = (0.0, 0.0, 1.0)
    # a fan of 40 faces around vertex c
    c = vert((0, 0, 0), up, (0.5, 0.5))
    ring = [vert((np.cos(a), np.sin(a), 0.1 * np.sin(3 * a)), _unit([0.1 * np.cos(a), 0.1 * np.sin(a), 1.0]),
                 (0.5 + 0.5 * np.cos(a), 0.5 + 0.5 * np.sin(a))) for a in np.linspace(0, 2 * np.pi, 41)[:-1]]
    for k in range(40):
        I += [c, ring[k], ring[(k + 1) % 40]]
    # a butterfly: four faces on the edge (a, b)
    a = vert((3, 0, 0), up, (0, 0))
    b = vert((3, 1, 0), up, (0, 1))
    for k, q in enumerate([(4, 0.5, 0), (2, 0.5, 0), (3.5, 0.5, 1), (2.5, 0.5, -1)]):
        w = vert(q, _unit([0.2 * k, 0.1, 1]), (1 if k % 2 == 0 else -1, 0.5))
        I += [a, b, w] if k % 2 == 0 else [b, a, w]
    # a duplicate of a and b with equal attributes, used by one more face
    a2 = vert((3, 0, 0), up, (0, 0))
    b2 = vert((3, 1, 0), up, (0, 1))
    I += [b2, a2, vert((3.2, 0.5, 0.3), up, (0.4, 0.6))]
    # zero uv area: all three uvs equal (isolated), and collinear uvs next to a good face
    I += [vert((6, 0, 0), up, (0.3, 0.3)), vert((7, 0, 0), up, (0.3, 0.3)), vert((6, 1, 0), up, (0.3, 0.3))]
    p = vert((8, 0, 0), up, (0, 0))
    q = vert((9, 0, 0), up, (1, 0))
    r = vert((8, 1, 0), up, (2, 0))
    s = vert((9, 1, 0), up, (1, 1))
    I += [p, q, r, q, s, r]
    # degenerate positions: two equal corners; a face of one repeated vertex index
    d = vert((10, 0, 0), up, (0, 0))
    I += [d, vert((10, 0, 0), up, (1, 0)), vert((11, 1, 0), up, (1, 1))]
    I += [c, c, ring[3]]
    # a mirrored face next to a normal one (orientations differ across the edge)
    m0 = vert((12, 0, 0), up, (0, 0))
    m1 = vert((13, 0, 0), up, (1, 0))
    m2 = vert((12, 1, 0), up, (0, 1))
    m3 = vert((13, 1, 0), up, (-1, 1))
    I += [m0, m1, m2, m1, m3, m2]
    # a zero-uv-area face after a good face of positive uv orientation across an
    # edge: it takes that group's orientation and frames
    g0, g1, g2 = vert((14, 0, 0), up, (0, 0)), vert((15, 0, 0), up, (1, 0)), vert((14, 1, 0), up, (0, 1))
    I += [g0, g1, g2, g2, g1, vert((15, 1, 0), up, (2, -1))]
    # a zero-uv-area face between groups of opposite orientation: the first
    # group in face order to reach it sets its orientation
    k0, k1, k2 = vert((16, 0, 0), up, (0, 0)), vert((17, 0, 0), up, (0, 1)), vert((16, 1, 0), up, (1, 0))
    k3, k4 = vert((17, 1, 0), up, (2, -1)), vert((18, 0.5, 0), up, (1, -3))
    I += [k3, k1, k4, k0, k1, k2, k2, k1, k3]
    # degenerate faces at vertices where groups of both orientations meet: they
    # copy the frame of the first good corner of the vertex
    I += [m1, m1, m2, m2, m1, m2]
    return (np.asarray(P, np.float32), _unit(N), np.asarray(U, np.float32), np.asarray(I, np.uint32))


def _soup(nv, nf, seed):
    """Random faces over a small vertex pool (many shared and non-manifold
    edges, mixed orientations, some repeated indices)."""
    rng = np.random.default_rng(seed)
    P = rng.uniform(-1, 1, (nv, 3)).astype(np.float32)
    N = _unit(rng.normal(size=(nv, 3)))
    U = rng.uniform(0, 1, (nv, 2)).astype(np.float32)
    idx = rng.integers(0, nv, (nf, 3)).astype(np.uint32)
    return P, N, U, idx.reshape(-1)


MESHES = {
    "grid": lambda: _height_grid(24, 1),
    "mirrored_grid": lambda: _height_grid(24, 2, mirror=True),
    "flat_grid": lambda: _height_grid(8, 3, bumps=0.0),
    "sphere": lambda: _sphere(32, 16),
    "awkward": _awkward,
    "soup": lambda: _soup(60, 400, 4),
    "soup_dense": lambda: _soup(12, 300, 5),
}


@pytest.mark.parametrize("name", list(MESHES))
def test_tangents_equal_reference_mikktspace(ref, name):
    P, N, U, idx = MESHES[name]()
    got = ptcore.generate_tangents(P, N, U, idx)
    want = ref(P, N, U, idx)
    assert got.shape == want.shape
    bad = ~np.all(got == want, axis=1)
    assert not bad.any(), (name, np.flatnonzero(bad)[:8], got[bad][:4], want[bad][:4])


def test_tangents_closed_form_on_a_plane():
    """A flat grid (normal +z) under four uv maps: the tangent is the unit
    dP/du and sign * (n x t) is the unit dP/dv."""
    P, N, U, idx = _height_grid(6, 0, bumps=0.0)
    for su, sv in [(1, 1), (-1, 1), (1, -1), (-1, -1)]:
        t = ptcore.generate_tangents(P, N, U * np.float32([su, sv]), idx)
        assert np.allclose(t[:, :3], [su, 0, 0], atol=1e-6), (su, sv)
        bitangent = t[:, 3:4] * np.cross(N, t[:, :3])
        assert np.allclose(bitangent, [0, sv, 0], atol=1e-6), (su, sv)


def test_tangents_unit_and_orthogonal_to_the_normal():
    P, N, U, idx = _sphere(24, 12)
    t = ptcore.generate_tangents(P, N, U, idx)
    used = np.zeros(len(P), bool)
    used[idx] = True
    assert np.allclose(np.linalg.norm(t[used, :3], axis=1), 1.0, atol=1e-5)
    assert np.abs(np.sum(t[used, :3] * N[used], axis=1)).max() < 1e-5
    # away from the poles and the u seam (where only one side's faces count)
    # the angle-weighted mean of the chords is d/du of the sphere, (-sin phi, cos phi, 0)
    ph = U[:, 0] * 2 * np.pi
    mid = used & (np.abs(P[:, 2]) < 0.9) & (U[:, 0] > 0) & (U[:, 0] < 1)
    assert np.allclose(t[mid, :3], np.stack([-np.sin(ph), np.cos(ph), 0 * ph], 1)[mid], atol=0.05)


def test_tangents_edge_cases(ref):
    P, N, U, idx = _height_grid(2, 0)
    # no faces: every vertex zero
    assert not ptcore.generate_tangents(P, N, U, idx[:0]).any()
    # an unused vertex stays zero, as in the reference
    extra = np.concatenate([P, [[5, 5, 5]]]).astype(np.float32)
    t = ptcore.generate_tangents(extra, np.concatenate([N, [[0, 0, 1]]]), np.concatenate([U, [[0, 0]]]), idx)
    assert not t[-1].any()
    with pytest.raises(ptcore.PtError):
        ptcore.generate_tangents(P, N, U, idx[:4])
    with pytest.raises(ptcore.PtError):
        ptcore.generate_tangents(P, N, U, np.array([0, 1, len(P)], np.uint32))
