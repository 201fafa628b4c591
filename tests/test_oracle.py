"""Pins the CPU oracle against published vectors, closed forms, brute force and
the mathematics of each stage.  CPU only."""
import math

import numpy as np
import pytest

import scenes


# ---------------------------------------------------------------- PCG32

def test_pcg32_known_answer(oracle_lib):
    # pcg32-demo of pcg-c-basic (O'Neill), pcg32_srandom_r(42u, 54u): first six outputs
    got = oracle_lib.pcg32(42, 54, 6)
    assert [int(x) for x in got] == [0xa15c02b7, 0x7b47f409, 0xba1d3330, 0x83d2f293, 0xbfa4784b, 0xcbed606e]


@pytest.mark.parametrize("k", [0, 1, 2, 7, 100, 12345])
def test_pcg32_advance_equals_stepping(oracle_lib, k):
    seq = oracle_lib.pcg32(1, 977, k + 5)
    jumped = oracle_lib.pcg32(1, 977, 5, advance=k)
    assert np.array_equal(seq[k:], jumped)


def test_pcg32_streams_differ_and_are_uniform(oracle_lib):
    a = oracle_lib.pcg32(1, 0, 100000)
    b = oracle_lib.pcg32(1, 1, 100000)
    assert not np.array_equal(a, b)
    u = (a >> 8).astype(np.float64) / 2 ** 24
    hist = np.histogram(u, bins=16, range=(0, 1))[0]
    chi2 = ((hist - len(u) / 16) ** 2 / (len(u) / 16)).sum()
    assert chi2 < 40.0   # 15 dof, p ~ 5e-4


# ---------------------------------------------------------------- triangle test / traversal

def _mt_float64(rays, P):
    """Ray-triangle intersection in float64 by barycentric solve (Cramer's rule)."""
    o, d = rays[:, 0:3].astype(np.float64), rays[:, 4:7].astype(np.float64)
    v0, v1, v2 = P[:, 0].astype(np.float64), P[:, 1].astype(np.float64), P[:, 2].astype(np.float64)
    e1, e2 = v1 - v0, v2 - v0
    M = np.stack([-d, e1, e2], axis=2)
    rhs = o - v0
    sol = np.linalg.solve(M, rhs[..., None])[..., 0]
    return sol[:, 0], sol[:, 1], sol[:, 2]


def test_closest_hit_matches_float64_geometry(cornell, cornell_oracle):
    rays = scenes.random_rays(20000, (-1, 0, -1), (1, 2, 1), 11)
    hits = cornell_oracle.closest(rays, brute=True)
    prim = hits[:, 0]
    hit = prim != 0xFFFFFFFF
    assert hit.mean() > 0.5
    t, u, v = (hits[:, 1:].view(np.float32)[:, k].astype(np.float64) for k in range(3))
    t64, u64, v64 = _mt_float64(rays[hit], cornell.positions[prim[hit]])
    assert np.allclose(t[hit], t64, rtol=1e-4, atol=1e-5)
    assert np.allclose(u[hit], u64, atol=1e-4)
    assert np.allclose(v[hit], v64, atol=1e-4)
    assert (t[hit] >= 1e-4).all() and (u[hit] >= 0).all() and (v[hit] >= 0).all() and (u[hit] + v[hit] <= 1 + 1e-6).all()


def test_closest_hit_is_nearest_of_all_float64_hits(cornell, cornell_oracle):
    """A planted wrong sign/index would pick a farther triangle: compare with
    the float64 minimum over every triangle."""
    rays = scenes.random_rays(2000, (-1, 0, -1), (1, 2, 1), 12)
    hits = cornell_oracle.closest(rays, brute=True)
    T = cornell.triangle_count
    best = np.full(len(rays), np.inf)
    for tri in range(T):
        P = np.repeat(cornell.positions[tri:tri + 1], len(rays), axis=0)
        with np.errstate(all="ignore"):
            t, b1, b2 = _mt_float64(rays, P)
        ok = (b1 >= -1e-6) & (b2 >= -1e-6) & (b1 + b2 <= 1 + 1e-6) & (t >= 1e-4)
        best = np.where(ok & (t < best), t, best)
    t = hits[:, 1].view(np.float32).astype(np.float64)
    hit = hits[:, 0] != 0xFFFFFFFF
    assert np.array_equal(hit, np.isfinite(best))
    assert np.allclose(t[hit], best[hit], rtol=1e-4, atol=1e-5)


def test_reference_traversal_equals_brute_force_except_ties(cornell_oracle):
    rays = scenes.random_rays(200000, (-1, 0, -1), (1, 2, 1), 5)
    ref = cornell_oracle.closest(rays)
    brute = cornell_oracle.closest(rays, brute=True)
    diff = (ref != brute).any(axis=1)
    # every disagreement is an exact tie in t between coplanar triangles
    assert np.array_equal(ref[diff, 1], brute[diff, 1])
    assert diff.mean() < 0.01


def test_brute_force_tie_rule_on_degenerate_geometry(oracle_lib):
    """Duplicated, coplanar and zero-area triangles: the reference-order
    traversal and the brute force differ only on exact ties in t, where the
    brute force keeps the lowest primitive id; zero-area triangles never hit."""
    import test_degenerate as D
    s = D._torture_scene()
    o = oracle_lib.OracleScene(s)
    rays = D._rays(1 << 15, 11)
    ref = o.closest(rays)
    brute = o.closest(rays, brute=True)
    diff = (ref != brute).any(axis=1)
    assert diff.sum() > 100                                  # the scene is full of ties
    assert np.array_equal(ref[diff, 1], brute[diff, 1])      # same t
    assert (brute[diff, 0] < ref[diff, 0]).all()             # lower id wins
    hit = brute[:, 0] != 0xFFFFFFFF
    assert not np.isin(brute[hit, 0], list(D.ZERO_AREA)).any()
    # the duplicates of grid triangles 0-39 never beat their originals
    assert not np.isin(brute[hit, 0], list(D.DUPLICATES)).any()


def test_any_hit_equals_closest_below_tmax(cornell, cornell_oracle):
    rays = scenes.random_rays(50000, (-1, 0, -1), (1, 2, 1), 6)
    rng = np.random.default_rng(0)
    rays[:, 7] = rng.uniform(0.01, 3.0, len(rays)).astype(np.float32)
    occ = cornell_oracle.any(rays)
    occ_b = cornell_oracle.any(rays, brute=True)
    hits = cornell_oracle.closest(rays, brute=True)
    expect = (hits[:, 0] != 0xFFFFFFFF).astype(np.uint32)
    assert np.array_equal(occ, expect)
    assert np.array_equal(occ_b, expect)
    sh = scenes.shadow_rays(cornell, 20000, 3)
    assert np.array_equal(cornell_oracle.any(sh), cornell_oracle.any(sh, brute=True))


def test_reference_bvh_invariants(oracle_lib):
    s = scenes.spheres_scene(count=3, subdivisions=3, lights=16, seed=4)
    o = oracle_lib.OracleScene(s)
    nodes, order = o.bvh()
    T = s.triangle_count
    assert sorted(order.tolist()) == list(range(T))
    leaf_tris = np.zeros(T, np.int32)
    lo, hi = nodes[:, 0:3], nodes[:, 3:6]
    ints = nodes[:, 6:10].view(np.uint32)
    max_depth = min(64, math.ceil(2 * math.log2(T)))
    assert o.info["max_depth"] == max_depth
    stack = [(0, 0)]
    P = s.positions[order]
    while stack:
        i, depth = stack.pop()
        left, right, first, count = ints[i]
        if left == 0xFFFFFFFF:
            leaf_tris[first:first + count] += 1
            assert count <= o.info["leaf_size"] or depth >= max_depth
            pts = P[first:first + count].reshape(-1, 3)
            assert (pts >= lo[i] - 1e-6).all() and (pts <= hi[i] + 1e-6).all()
        else:
            for c in (left, right):
                assert (lo[c] >= lo[i] - 1e-4).all() and (hi[c] <= hi[i] + 1e-4).all()
                stack.append((c, depth + 1))
    assert (leaf_tris == 1).all()


# ---------------------------------------------------------------- alias table

def _alias_mass(bins):
    L = len(bins)
    mass = bins["q"].astype(np.float64).copy()
    for j in range(L):
        if bins["alias"][j] != 0xFFFFFFFF:
            mass[bins["alias"][j]] += 1.0 - bins["q"][j]
    return mass / L


@pytest.mark.parametrize("seed,L", [(0, 2), (1, 7), (2, 1024), (3, 33)])
def test_alias_table_reconstructs_distribution(oracle_lib, seed, L):
    rng = np.random.default_rng(seed)
    w = rng.lognormal(0, 1, L).astype(np.float32)
    bins = oracle_lib.alias_build(w)
    p = w.astype(np.float64) / w.astype(np.float64).sum()
    assert np.allclose(bins["p"], p, rtol=1e-6)
    assert np.allclose(_alias_mass(bins), p, atol=4 * L * 1.2e-7 + 1e-7)


def test_alias_sampling_frequencies(oracle_lib):
    rng = np.random.default_rng(7)
    w = rng.lognormal(0, 1, 64).astype(np.float32)
    bins = oracle_lib.alias_build(w)
    u = rng.random(1_000_000).astype(np.float32)
    idx, pmf = oracle_lib.alias_sample(bins, u)
    p = w / w.sum()
    counts = np.bincount(idx, minlength=64)
    expected = p * len(u)
    chi2 = ((counts - expected) ** 2 / expected).sum()
    assert chi2 < 120   # 63 dof
    assert np.array_equal(pmf, bins["p"][idx])


def test_alias_single_bin_and_edges(oracle_lib):
    bins = oracle_lib.alias_build(np.array([3.0], np.float32))
    idx, pmf = oracle_lib.alias_sample(bins, np.array([0.0, 0.5, 0.99999994], np.float32))
    assert idx.tolist() == [0, 0, 0] and np.allclose(pmf, 1.0)
    bins = oracle_lib.alias_build(np.array([1.0, 1.0, 2.0], np.float32))
    # u = 1 - 2^-24 must stay inside the last bin (AliasTableView.hpp:37-38 clamps)
    idx, _ = oracle_lib.alias_sample(bins, np.array([0.99999994], np.float32))
    assert idx[0] < 3


# ---------------------------------------------------------------- camera and tonemap

def test_camera_basis_closed_form(oracle_lib):
    view = scenes.View(fov_y=math.radians(40), position=(0.0, 1.0, 3.8), rotation=(0.0, 0.0, 0.0, 1.0))
    b = oracle_lib.camera_basis(view, 1.0).reshape(4, 3)
    half = math.tan(math.radians(20))
    origin, llc, hor, ver = b
    assert np.allclose(origin, [0, 1, 3.8])
    assert np.allclose(hor, [2 * half, 0, 0], atol=1e-6)
    assert np.allclose(ver, [0, 2 * half, 0], atol=1e-6)
    # image centre maps to forward
    assert np.allclose(llc + 0.5 * hor + 0.5 * ver - origin, [0, 0, -1], atol=1e-6)
    # a rotated view keeps an orthonormal basis with the look direction
    q = scenes.look_at_quaternion((1, 2, 3), (4, 0, -2))
    b = oracle_lib.camera_basis(scenes.View(0.8, (1, 2, 3), q), 16 / 9).reshape(4, 3)
    fwd = b[1] + 0.5 * b[2] + 0.5 * b[3] - b[0]
    want = np.array([3, -2, -5], np.float64)
    assert np.allclose(fwd / np.linalg.norm(fwd), want / np.linalg.norm(want), atol=1e-5)
    assert abs(np.dot(b[2], b[3])) < 1e-5


def test_tonemap_exact_values(oracle_lib):
    hdr = np.zeros((6, 4), np.float32)
    hdr[:, 0] = [0.0, 1.0, 1e9, 0.25, 3.0, 0.01]
    out = oracle_lib.tonemap(hdr)
    r = out & 0xFF
    expect = [int(min(max((c / (c + 1)) ** (1 / 2.2) * 255, 0), 255)) for c in hdr[:, 0].astype(np.float64)]
    assert (np.abs(r.astype(int) - np.array(expect)) <= 1).all()
    assert r[0] == 0 and r[2] == 255 and (out >> 24 == 255).all()
    assert r[1] == int(0.5 ** (1 / 2.2) * 255)


# ---------------------------------------------------------------- shading, closed forms

def test_direct_view_of_the_emitter_reads_le(oracle_lib):
    """Depth 1 without NEE traces camera rays only: a pixel reads Le times the
    fraction of its samples that hit the emitter (EvaluateEmissiveKernel,
    MaterialEval.cu:882-912), so every pixel is a multiple of Le / spp.  The
    closed forms of diffuse, GGX and mirror transport are in
    test_shading_closed_forms.py."""
    s = scenes.cornell_box()
    o = oracle_lib.OracleScene(s)
    settings = dict(width=8, height=8, samples_per_pixel=4, path_depth=1, use_nee=0, rr_start_depth=100, seed=1)
    hdr1, ext1, sh1 = o.render(settings, s.view)
    assert sh1 == 0 and ext1 == 8 * 8 * 4   # depth 1: camera rays only, no NEE
    assert (hdr1[:, 0] > 0).any()
    assert np.allclose(hdr1[:, :3] * 4 % 15, 0, atol=1e-3)


def test_nee_and_bsdf_sampling_agree(cornell, cornell_oracle):
    """NEE on and NEE off are two estimators of the same image
    (MaterialEval.cu:897-904): their means agree within 3 sigma."""
    base = dict(width=16, height=16, samples_per_pixel=256, path_depth=4, rr_start_depth=100, seed=3)
    on, _, sh = cornell_oracle.render(dict(base, use_nee=1), cornell.view)
    off, _, sh0 = cornell_oracle.render(dict(base, use_nee=0, seed=9), cornell.view)
    assert sh > 0 and sh0 == 0
    m_on, m_off = on[:, :3].mean(), off[:, :3].mean()
    # per-pixel estimates are independent: standard error from pixel spread
    se = math.sqrt(on[:, :3].mean(1).var() / len(on) + off[:, :3].mean(1).var() / len(off))
    assert abs(m_on - m_off) < 3 * max(se, 1e-3) + 0.02 * m_off


def test_russian_roulette_keeps_mean(cornell, cornell_oracle):
    base = dict(width=16, height=16, samples_per_pixel=256, path_depth=6, use_nee=1, seed=5)
    rr, ext_rr, _ = cornell_oracle.render(dict(base, rr_start_depth=1), cornell.view)
    no, ext_no, _ = cornell_oracle.render(dict(base, rr_start_depth=100, seed=6), cornell.view)
    assert ext_rr < ext_no
    se = math.sqrt(rr[:, :3].mean(1).var() / len(rr) + no[:, :3].mean(1).var() / len(no))
    assert abs(rr[:, :3].mean() - no[:, :3].mean()) < 3 * max(se, 1e-3) + 0.02 * no[:, :3].mean()


def test_render_is_deterministic_and_thread_independent(cornell, cornell_oracle):
    st = dict(width=16, height=16, samples_per_pixel=16, path_depth=8, use_nee=1, rr_start_depth=3, seed=1)
    a, ea, sa = cornell_oracle.render(st, cornell.view, threads=1)
    b, eb, sb = cornell_oracle.render(st, cornell.view, threads=8)
    assert np.array_equal(a, b) and ea == eb and sa == sb
    c, _, _ = cornell_oracle.render(st, cornell.view, pixel_begin=40, pixel_end=80)
    assert np.array_equal(c, a[40:80])


# ---------------------------------------------------------------- golden fixtures

def test_oracle_reproduces_golden_fixtures(cornell, cornell_oracle, oracle_lib):
    import os
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "c1_hits_4096.npz"))
    assert np.array_equal(cornell_oracle.closest(g["rays"], brute=True), g["hits"])
    assert np.array_equal(cornell_oracle.classify(g["rays"]), g["queue"])
    assert np.array_equal(cornell_oracle.any(g["shadow_rays"]), g["occluded"])
    a = np.load(os.path.join(os.path.dirname(__file__), "golden", "alias_1024.npz"))
    bins = oracle_lib.alias_build(a["weights"])
    assert np.array_equal(bins.view(np.uint32).reshape(-1, 4), a["bins"])
    idx, pmf = oracle_lib.alias_sample(bins, a["u"])
    assert np.array_equal(idx, a["index"]) and np.array_equal(pmf, a["pmf"])
    r = np.load(os.path.join(os.path.dirname(__file__), "golden", "c1_render_16x16_64spp.npz"))
    st = dict(width=16, height=16, samples_per_pixel=64, path_depth=8, use_nee=1, rr_start_depth=3, seed=1)
    hdr, ext, sh = cornell_oracle.render(st, cornell.view)
    assert np.array_equal(hdr, r["hdr"]) and ext == r["ext"] and sh == r["shadow"]
