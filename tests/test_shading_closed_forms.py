"""Pins the oracle's shading (MaterialEval.cu:222-392,587-880) by mathematics.

Every check is one that a plausible slip fails: a missing pi or 4 cos cos, a
wrong power of alpha or cos, a one-sided Smith term, a density that forgets the
Jacobian 1/(4 wo.wm) of the reflection, a wrong lobe probability, a reflect
with the wrong sign, a throughput applied twice.  CPU only.
"""
import math
import zlib

import numpy as np
import pytest

import closed_form_scenes as cf
import scenes



@pytest.fixture
def rng(request):
    """A generator seeded by the test's own name: each case draws the same
    samples whatever runs before it."""
    return np.random.default_rng([20261204, zlib.crc32(request.node.name.encode())])


def _hemi(n, rng, min_z=0.02):
    """Random unit directions with z >= min_z."""
    z = rng.uniform(min_z, 1.0, n)
    phi = rng.uniform(0, 2 * math.pi, n)
    r = np.sqrt(1 - z * z)
    return np.stack([r * np.cos(phi), r * np.sin(phi), z], 1)


def _dir(theta, phi=0.3):
    return np.array([math.sin(theta) * math.cos(phi), math.sin(theta) * math.sin(phi), math.cos(theta)])


# ---------------------------------------------------------------- sampling trig

def test_sincos_of_sampling_angles_is_accurate(oracle_lib):
    """The shared sin/cos of 2 pi u (every u = k 2^-24 on a stride, and the
    quadrant edges) is within 1.6 ulp / 1.5e-7 of float64 sin/cos."""
    u = (np.arange(0, 1 << 24, 7, dtype=np.float64) * 2.0 ** -24)
    x = (2.0 * np.float32(math.pi) * u.astype(np.float32)).astype(np.float32)
    x = np.concatenate([x, np.float32([0.0, math.pi / 4, math.pi / 2, math.pi, 1.5 * math.pi, 6.2831850])])
    s, c = oracle_lib.sincos(x)
    es, ec = np.sin(x.astype(np.float64)), np.cos(x.astype(np.float64))
    assert np.abs(s - es).max() < 1.5e-7 and np.abs(c - ec).max() < 1.5e-7
    big = np.abs(es) > 1e-3
    assert (np.abs(s - es)[big] / np.spacing(np.abs(es[big]).astype(np.float32))).max() < 1.6
    big = np.abs(ec) > 1e-3
    assert (np.abs(c - ec)[big] / np.spacing(np.abs(ec[big]).astype(np.float32))).max() < 1.6


# ---------------------------------------------------------------- diffuse sampling

def test_cosine_hemisphere_density(rng, oracle_lib):
    """SampleCosineWeightedHemisphere (MaterialEval.cu:222-229): unit vectors
    above the surface whose density is cos(theta)/pi: z^2 is uniform and phi is
    uniform (chi-square over a 16 x 16 grid)."""
    n = 1 << 20
    u = rng.random((n, 2)).astype(np.float32)
    w = oracle_lib.cosine_hemisphere(u).astype(np.float64)
    assert np.allclose(np.linalg.norm(w, axis=1), 1.0, atol=1e-6) and (w[:, 2] >= 0).all()
    a = np.clip(w[:, 2] ** 2, 0, 1 - 1e-12)
    b = (np.arctan2(w[:, 1], w[:, 0]) % (2 * math.pi)) / (2 * math.pi)
    h = np.histogram2d(a, b, bins=16, range=[[0, 1], [0, 1]])[0]
    e = n / 256
    chi2 = ((h - e) ** 2 / e).sum()
    assert chi2 < 255 + 6 * math.sqrt(2 * 255)


# ---------------------------------------------------------------- GGX microfacet terms

@pytest.mark.parametrize("alpha", [0.05, 0.3, 0.8])
def test_ggx_d_is_normalised(oracle_lib, alpha):
    """integral of D(wm) cos(theta_m) over the hemisphere = 1 (Walter et al.
    2007, eq. 4), by the midpoint rule in psi with tan(theta) = alpha tan(psi)."""
    n = 200000
    psi = (np.arange(n) + 0.5) / n * (math.pi / 2)
    theta = np.arctan(alpha * np.tan(psi))
    dtheta = alpha / (np.cos(psi) ** 2 + (alpha * np.sin(psi)) ** 2)   # d theta / d psi
    wm = np.stack([np.sin(theta), np.zeros(n), np.cos(theta)], 1)
    D = oracle_lib.ggx_term("D", alpha, wm).astype(np.float64)
    integral = 2 * math.pi * (D * np.cos(theta) * np.sin(theta) * dtheta).sum() * (math.pi / 2 / n)
    assert abs(integral - 1.0) < 2e-4, integral


def test_ggx_terms_equal_published_forms(rng, oracle_lib):
    """D = a^2 / (pi cos^4 (a^2 + tan^2)^2), Lambda = (sqrt(1 + a^2 tan^2) - 1)/2,
    G1 = 1 / (1 + Lambda) at random directions and azimuths."""
    w = _hemi(4096, rng, min_z=0.05)
    for alpha in (0.02, 0.2, 0.6, 1.0):
        t2 = (1 - w[:, 2] ** 2) / w[:, 2] ** 2
        D = alpha ** 2 / (math.pi * w[:, 2] ** 4 * (alpha ** 2 + t2) ** 2)
        lam = (np.sqrt(1 + alpha ** 2 * t2) - 1) / 2
        assert np.allclose(oracle_lib.ggx_term("D", alpha, w), D, rtol=2e-5)
        assert np.allclose(oracle_lib.ggx_term("lambda", alpha, w), lam, rtol=2e-5, atol=1e-7)
        assert np.allclose(oracle_lib.ggx_term("G1", alpha, w), 1 / (1 + lam), rtol=2e-5)


def g1_d_abs(alpha, wo, wm):
    """G1(wo) |wo.wm| D(wm) / cos(theta_o), the reference's form of D_wo."""
    c = wm[:, 2]
    t2 = (1 - c * c) / (c * c)
    D = alpha ** 2 / (math.pi * c ** 4 * (alpha ** 2 + t2) ** 2)
    t2o = (1 - wo[2] ** 2) / wo[2] ** 2
    g1 = 1 / (1 + (math.sqrt(1 + alpha ** 2 * t2o) - 1) / 2)
    return g1 * np.abs(wm @ wo) * D / wo[2]


def _dvis64(alpha, wo, wm):
    """Visible-normal density D_wo(wm) = G1(wo) max(0, wo.wm) D(wm) / cos(theta_o)
    (Heitz 2018, eq. 1), float64."""
    c = wm[:, 2]
    t2 = (1 - c * c) / (c * c)
    D = alpha ** 2 / (math.pi * c ** 4 * (alpha ** 2 + t2) ** 2)
    t2o = (1 - wo[2] ** 2) / wo[2] ** 2
    g1 = 1 / (1 + (math.sqrt(1 + alpha ** 2 * t2o) - 1) / 2)
    return g1 * np.maximum(0.0, wm @ wo) * D / wo[2]


def _vndf64(wo, alpha, u, reference_basis):
    """Visible-normal sample of Heitz 2018 (JCGT 7(4), listing 1) in float64:
    stretch wo, basis T1 = (-Vh.y, Vh.x, 0)/|.|, T2 = Vh x T1, warp the
    uniform disk's second coordinate by s = (1 + Vh.z)/2, lift to the
    hemisphere, unstretch.  reference_basis: T1 = Vh x (0, 0, 1) instead, as
    TrowbridgeSampleWm builds it (MaterialEval.cu:600), which is -T1 (and
    T2 becomes -T2); its 0.999 threshold for the special case as well."""
    vh = np.array([alpha * wo[0], alpha * wo[1], wo[2]])
    vh /= np.linalg.norm(vh)
    if reference_basis:
        t1 = np.cross(vh, [0.0, 0.0, 1.0]) if vh[2] < 0.999 else np.array([1.0, 0.0, 0.0])
    else:
        l2 = vh[0] ** 2 + vh[1] ** 2
        t1 = np.array([-vh[1], vh[0], 0.0]) if l2 > 0 else np.array([1.0, 0.0, 0.0])
    t1 = t1 / np.linalg.norm(t1)
    t2 = np.cross(vh, t1)
    r = np.sqrt(u[:, 0])
    phi = 2 * math.pi * u[:, 1]
    a, b = r * np.cos(phi), r * np.sin(phi)
    h = np.sqrt(1 - a * a)
    s = 0.5 * (1 + vh[2])
    b = h + s * (b - h)
    nh = a[:, None] * t1 + b[:, None] * t2 + np.sqrt(np.maximum(0, 1 - a * a - b * b))[:, None] * vh
    wm = np.stack([alpha * nh[:, 0], alpha * nh[:, 1], np.maximum(1e-6, nh[:, 2])], 1)
    return wm / np.linalg.norm(wm, axis=1, keepdims=True)


def _vndf_chi2(alpha, wo, wm):
    """Chi-square of samples wm against D_wo over 24 x 24 bins in (psi, phi),
    tan(theta_m) = alpha tan(psi), the bins' masses by a sub-grid midpoint
    rule; also returns the total mass of D_wo (1 by construction)."""
    n = len(wm)
    psi = np.arctan(np.sqrt(wm[:, 0] ** 2 + wm[:, 1] ** 2) / wm[:, 2] / alpha)
    phi = np.arctan2(wm[:, 1], wm[:, 0]) % (2 * math.pi)
    nb, sub = 24, 24
    h = np.histogram2d(psi, phi, bins=[nb, nb], range=[[0, math.pi / 2], [0, 2 * math.pi]])[0]
    gp = (np.arange(nb * sub) + 0.5) / (nb * sub) * (math.pi / 2)
    gf = (np.arange(nb * sub) + 0.5) / (nb * sub) * (2 * math.pi)
    P, F = np.meshgrid(gp, gf, indexing="ij")
    th = np.arctan(alpha * np.tan(P))
    jac = alpha / (np.cos(P) ** 2 + (alpha * np.sin(P)) ** 2)
    w = np.stack([np.sin(th) * np.cos(F), np.sin(th) * np.sin(F), np.cos(th)], -1).reshape(-1, 3)
    dens = _dvis64(alpha, wo, w).reshape(P.shape) * np.sin(th) * jac
    cell = (math.pi / 2 / (nb * sub)) * (2 * math.pi / (nb * sub))
    mass = (dens * cell).reshape(nb, sub, nb, sub).sum((1, 3))
    e = mass * n
    keep = e > 20
    chi2 = ((h[keep] - e[keep]) ** 2 / e[keep]).sum() + (h[~keep].sum() - e[~keep].sum()) ** 2 / max(e[~keep].sum(), 1)
    return chi2, int(keep.sum()), mass.sum()


VNDF_CASES = [(0.1, 0.2), (0.5, 0.9), (0.9, 1.35), (0.3, 0.0)]


@pytest.mark.parametrize("alpha, theta_o", VNDF_CASES)
def test_published_vndf_method_samples_the_visible_normal_density(rng, alpha, theta_o):
    """Heitz's method (float64 restatement) draws wm with density
    D_wo(wm) = G1(wo) max(0, wo.wm) D(wm) / cos(theta_o): chi-square of 2^20
    samples; D_wo integrates to 1.  This pins the method the next test
    holds the oracle to."""
    wo = _dir(theta_o)
    u = rng.random((1 << 20, 2))
    chi2, dof, mass = _vndf_chi2(alpha, wo, _vndf64(wo, alpha, u, reference_basis=False))
    assert abs(mass - 1) < 2e-3, mass
    assert chi2 < dof + 6 * math.sqrt(2 * dof), (chi2, dof)


@pytest.mark.parametrize("alpha, theta_o", VNDF_CASES)
def test_oracle_vndf_is_the_published_method_in_the_reference_basis(rng, oracle_lib, alpha, theta_o):
    """TrowbridgeSampleWm (MaterialEval.cu:594-610), as the oracle restates it,
    equals Heitz's method sample for sample with the tangent basis the
    reference builds, cross(wh, z) = -T1.  That basis moves the projected-area
    warp to the other half of the disk: the reference's samples follow D_wo
    only at normal incidence (where the warp vanishes) and drift from it as
    the view tilts, a bias of the reference that the core reproduces."""
    wo = _dir(theta_o)
    n = 1 << 20
    u = rng.random((n, 2)).astype(np.float32)
    wm = oracle_lib.ggx_sample_wm(alpha, np.broadcast_to(wo, (n, 3)), u).astype(np.float64)
    ref = _vndf64(wo, alpha, u.astype(np.float64), reference_basis=True)
    # float32 against float64: the basis is ill-conditioned for the rare
    # half-vectors near the pole, every other sample agrees to 2e-5
    assert np.abs(wm - ref).max() < 2e-3
    assert np.quantile(np.abs(wm - ref).max(1), 0.999) < 2e-5
    chi2, dof, _ = _vndf_chi2(alpha, wo, wm)
    if theta_o == 0.0:
        assert chi2 < dof + 6 * math.sqrt(2 * dof), (chi2, dof)
    else:   # the reference's bias is measurable at every tilted view checked here
        assert chi2 > dof + 12 * math.sqrt(2 * dof), (chi2, dof)
    # the oracle's D_wo is the published density on the visible side (wo.wm > 0,
    # the only side the estimator evaluates); on the back side the reference
    # takes |wo.wm| where the published form has max(0, wo.wm) (MaterialEval.cu:642-645)
    w = _hemi(4096, rng, 0.05)
    got = oracle_lib.ggx_dvis(alpha, np.broadcast_to(wo, w.shape), w)
    vis = w @ wo > 0
    assert np.allclose(got[vis], _dvis64(alpha, wo, w[vis]), rtol=5e-5, atol=1e-6)
    assert np.allclose(got[~vis], g1_d_abs(alpha, wo, w[~vis]), rtol=5e-5, atol=1e-6)


# ---------------------------------------------------------------- GGX BSDF

GGX_MATERIALS = [
    ((1.0, 1.0, 1.0), 0.5, 1.0, 1.5),     # metal, F = 1
    ((0.8, 0.4, 0.2), 0.3, 0.5, 1.5),     # half metallic, coloured
    ((0.7, 0.7, 0.7), 0.9, 0.0, 1.5),     # rough dielectric
    ((0.2, 0.6, 0.9), 0.2, 0.0, 1.33),    # smooth dielectric
]


@pytest.mark.parametrize("material", GGX_MATERIALS)
def test_ggx_bsdf_is_reciprocal_and_equals_published_form(rng, oracle_lib, material):
    """f(wo, wi) = f(wi, wo) (Helmholtz), and f equals (1 - F) diffuse / pi +
    F D G / (4 cos_i cos_o) evaluated in float64 from the published forms."""
    n = 20000
    wo, wi = _hemi(n, rng, 0.05), _hemi(n, rng, 0.05)
    f, ok = oracle_lib.ggx_bsdf(material, wo, wi)
    fr, okr = oracle_lib.ggx_bsdf(material, wi, wo)
    assert ok.mean() > 0.99 and np.array_equal(ok, okr)
    assert np.allclose(f[ok], fr[ok], rtol=1e-5, atol=1e-7)
    ref = cf.ggx_f64(material, wo, wi)
    assert np.allclose(f[ok], ref[ok], rtol=2e-4, atol=1e-6)


def _albedo_quadrature(material, wo, n_c2=1500, n_phi=1500):
    """integral of f(wo, wi) cos(theta_i) over the hemisphere: midpoint rule in
    (cos^2 theta_i, phi), where cos d(omega) = d(cos^2) d(phi) / 2."""
    c2 = (np.arange(n_c2) + 0.5) / n_c2
    ph = (np.arange(n_phi) + 0.5) / n_phi * 2 * math.pi
    C2, PH = np.meshgrid(c2, ph, indexing="ij")
    ct, st = np.sqrt(C2).ravel(), np.sqrt(1 - C2).ravel()
    wi = np.stack([st * np.cos(PH.ravel()), st * np.sin(PH.ravel()), ct], 1)
    f = cf.ggx_f64(material, np.broadcast_to(wo, wi.shape), wi)
    return f.sum(0) * (1 / n_c2) * (2 * math.pi / n_phi) / 2


@pytest.mark.parametrize("material", GGX_MATERIALS)
@pytest.mark.parametrize("theta_o", [0.0, 0.3, 1.1])
def test_ggx_sampling_estimates_the_albedo_and_keeps_energy(rng, oracle_lib, material, theta_o):
    """EvaluateGgxIndirect (MaterialEval.cu:720-836): the mean of its weight
    f cos / pdf is the directional albedo, the integral of f cos (quadrature of
    the published BSDF), which never exceeds 1 (no emission: the GGX
    furnace).  Exact at normal incidence; at tilted views the reference's
    visible-normal basis (previous tests) biases it, by under 0.2 % at 17
    degrees and under 2 % at 63 degrees.  Each weight equals f cos / pdf of the
    evaluation functions."""
    wo = _dir(theta_o)
    n = 1 << 21
    u = rng.random((n, 3)).astype(np.float32)
    wl, f, wt, pdf, ok = oracle_lib.ggx_sample(material, np.broadcast_to(wo, (n, 3)), u)
    est = np.where(ok[:, None], wt, 0).astype(np.float64)
    mean, se = est.mean(0), est.std(0) / math.sqrt(n)
    ref = _albedo_quadrature(material, wo)
    bias = {0.0: 2e-4, 0.3: 2e-3, 1.1: 2e-2}[theta_o]
    assert (np.abs(mean - ref) < 4 * se + bias * ref).all(), (mean, ref, se)
    assert (ref <= 1.0 + 1e-6).all() and (mean <= 1.0 + 4 * se).all()
    # per sample: weight = f cos / pdf with f and pdf from the evaluation functions
    k = np.nonzero(ok)[0][:50000]
    fe, oke = oracle_lib.ggx_bsdf(material, np.broadcast_to(wo, (len(k), 3)), wl[k])
    pe = oracle_lib.ggx_pdf(material, np.broadcast_to(wo, (len(k), 3)), wl[k])
    assert np.allclose(pdf[k], pe, rtol=1e-6)
    m = oke
    assert m.mean() > 0.999
    assert np.allclose(f[k][m], fe[m], rtol=1e-6, atol=1e-9)
    assert np.allclose(wt[k], f[k] * (wl[k, 2:3] / pdf[k, None]), rtol=1e-5)


@pytest.mark.parametrize("material, theta_o", [(GGX_MATERIALS[1], 0.5), (GGX_MATERIALS[3], 1.2),
                                               (((1.0, 1.0, 1.0), 0.3, 1.0, 1.5), 0.7)])
def test_ggx_sampled_directions_follow_the_stated_density(rng, oracle_lib, material, theta_o):
    """The directions EvaluateGgxIndirect draws (diffuse or visible-normal lobe
    picked by luminance, then reflected) have the density the estimator
    divides by: pd cos/pi + ps D_wo(wm) / (4 wo.wm).  Chi-square over bins in
    (cos^2 theta_i, phi); the stated density is checked against the published
    forms in float64."""
    wo = _dir(theta_o, phi=0.0)
    n = 1 << 21
    u = rng.random((n, 3)).astype(np.float32)
    wl, _, _, pdf, ok = oracle_lib.ggx_sample(material, np.broadcast_to(wo, (n, 3)), u)
    a = np.clip(wl[ok, 2].astype(np.float64) ** 2, 0, 1 - 1e-12)
    b = np.arctan2(wl[ok, 1], wl[ok, 0]) % (2 * math.pi)
    nb, sub = 24, 16
    h = np.histogram2d(a, b, bins=nb, range=[[0, 1], [0, 2 * math.pi]])[0]
    g2 = (np.arange(nb * sub) + 0.5) / (nb * sub)
    gf = (np.arange(nb * sub) + 0.5) / (nb * sub) * 2 * math.pi
    C2, PH = np.meshgrid(g2, gf, indexing="ij")
    ct, st = np.sqrt(C2).ravel(), np.sqrt(1 - C2).ravel()
    w = np.stack([st * np.cos(PH.ravel()), st * np.sin(PH.ravel()), ct], 1)
    dens = oracle_lib.ggx_pdf(material, np.broadcast_to(wo, w.shape), w).astype(np.float64)
    # d(omega) = d(cos^2) d(phi) / (2 cos)
    mass = (dens / (2 * ct) * (1 / (nb * sub)) * (2 * math.pi / (nb * sub))).reshape(nb, sub, nb, sub).sum((1, 3))
    e = mass * n
    keep = e > 20
    chi2 = ((h[keep] - e[keep]) ** 2 / e[keep]).sum()
    dof = int(keep.sum())
    assert chi2 < dof + 6 * math.sqrt(2 * dof), (chi2, dof)
    assert abs(h[~keep].sum() - e[~keep].sum()) < 6 * math.sqrt(e[~keep].sum() + 10)
    # the stated density in float64 from the published forms
    color, rough, metal, ior = material
    alpha = max(rough * rough, 1e-3)
    base = np.asarray(color, np.float64)
    f0 = ((1 - ior) / (1 + ior)) ** 2 * (1 - metal) + base * metal
    lum = np.array([0.2126, 0.7152, 0.0722])
    dw, sw = lum @ (base * (1 - metal)), lum @ f0
    pd, ps = dw / (dw + sw), sw / (dw + sw)
    sel = slice(0, None, 97)
    wm = w[sel] + wo
    wm /= np.linalg.norm(wm, axis=1, keepdims=True)
    ref = pd * w[sel, 2] / math.pi + ps * _dvis64(alpha, wo, wm) / (4 * (wm @ wo))
    assert np.allclose(dens[sel], ref, rtol=2e-4, atol=1e-6)


def test_ggx_smooth_metal_reflects_all_energy(rng, oracle_lib):
    """At the smallest alpha (1e-3, MaterialEval.cu:860) a metal of base colour
    1 (F = 1) reflects wo about the normal with weight 1: G -> 1, and the
    visible-normal sample is the normal."""
    material = ((1.0, 1.0, 1.0), 0.0, 1.0, 1.5)
    wo = _dir(0.6, phi=1.0)
    n = 1 << 16
    u = rng.random((n, 3)).astype(np.float32)
    wl, _, wt, _, ok = oracle_lib.ggx_sample(material, np.broadcast_to(wo, (n, 3)), u)
    assert ok.mean() > 0.999
    mirror = np.array([-wo[0], -wo[1], wo[2]])
    # GGX keeps long tails at any alpha (P(tan theta_m > k alpha) = 1/(1 + k^2)):
    # half the samples within 4 alpha of the mirror direction, 99.9 % within 70 alpha
    dev = np.abs(wl[ok] - mirror).max(1)
    assert np.median(dev) < 4e-3 and np.quantile(dev, 0.999) < 0.07
    assert abs(wt[ok].mean() - 1.0) < 3e-3 and (wt[ok] <= 1.0 + 1e-3).all()


# ---------------------------------------------------------------- scene-level closed forms (oracle render)

def _block_means(hdr, w, h, blk):
    img = hdr[:, :3].mean(1).reshape(h, w)
    return img.reshape(h // blk, blk, w // blk, blk).mean((1, 3))


@pytest.mark.parametrize("nee, depth, rr", [(1, 8, 100), (0, 8, 100), (1, 1, 100), (0, 2, 1), (1, 8, 1)])
def test_diffuse_plates_match_form_factor(oracle_lib, nee, depth, rr):
    """Diffuse floor under a parallel emitter: E[L(p)] = rho Le F(p) with the
    form factor F of Howell's catalogue, for NEE on (light sampling,
    MaterialEval.cu:231-268) and off (cosine sampling, :270-309), with and
    without roulette; depth 1 without NEE sees only the floor (0)."""
    rho, le = 0.6, 3.0   # light samples stay below the 20-unit shadow clamp (at most 9.2)
    s = cf.plates(rho, le)
    o = oracle_lib.OracleScene(s)
    w = h = 32
    st = dict(width=w, height=h, samples_per_pixel=256, path_depth=depth, use_nee=nee, rr_start_depth=rr, seed=7)
    hdr, ext, sh = o.render(st, s.view)
    cf.check_plates(hdr, s.view, w, h, rho, le, nee, depth)
    if depth == 1 and not nee:
        assert sh == 0


FURNACE_CASES = [(1, 1, 100), (1, 3, 100), (1, 8, 100), (0, 2, 100), (0, 3, 100), (0, 8, 100), (1, 8, 1), (0, 8, 2)]


@pytest.mark.parametrize("nee, depth, rr", FURNACE_CASES)
def test_diffuse_sphere_furnace_is_a_geometric_series(oracle_lib, nee, depth, rr):
    """Inside a sphere with a quarter of its area emissive: each diffuse
    bounce reaches emissive area with probability f and carries rho, so a
    diffuse pixel reads rho f Le (1 + q + ... + q^(n-1)), q = rho (1 - f), for
    NEE on (n = depth, light sampling by area, MaterialEval.cu:231-268) and off
    (n = depth - 1, cosine sampling, :270-309), with and without roulette
    (:313-350).  A missing albedo, a wrong pdf, an emitter counted twice under
    NEE or a bounce more or less each move the mean by many standard errors."""
    rho, le = 0.8, 2.0
    s = cf.sphere_furnace(rho, le)
    f = cf.furnace_fraction(s)
    assert abs(f - 0.25) < 1e-6
    o = oracle_lib.OracleScene(s)
    w = h = 32
    closest = lambda r: o.closest(r, tprim=True)
    mask = cf.furnace_pixels(s, closest, w, h)
    assert mask.sum() > 600
    st = dict(width=w, height=h, samples_per_pixel=256, path_depth=depth, use_nee=nee, rr_start_depth=rr, seed=4)
    hdr, ext, sh = o.render(st, s.view)
    cf.check_furnace(hdr, mask, rho, le, f, depth, nee)
    # emissive pixels read Le: a camera ray that meets an emitter counts it
    emit = cf.furnace_pixels(s, closest, w, h, material=1)
    assert emit.sum() > 50 and np.allclose(hdr[emit, :3], le, rtol=1e-5)


def test_mirror_corridor_is_a_power_series(oracle_lib):
    """Two parallel mirrors of reflectance s and an emitter across the end:
    a camera ray reflects k times (reflect(d, n) = d - 2 (n.d) n,
    MaterialEval.cu:352-392) and reads exactly s^k Le; rays needing k >= the
    depth limit read 0.  Without roulette every sample of a pixel is the same."""
    s = cf.corridor()
    o = oracle_lib.OracleScene(s)
    w, h = 64, 16
    for depth in (8, 4):
        st = dict(width=w, height=h, samples_per_pixel=16, path_depth=depth, use_nee=1, rr_start_depth=100, seed=3)
        hdr, _, sh = o.render(st, s.view)
        assert sh == 0
        cf.check_corridor(hdr, s.view, w, h, depth, 16, 100)


def test_mirror_corridor_roulette_keeps_the_mean(oracle_lib):
    """With roulette from the first bounce a path survives k reflections with
    probability prod q and carries s^k Le / prod q: the mean stays s^k Le."""
    s = cf.corridor()
    o = oracle_lib.OracleScene(s)
    w, h = 64, 16
    st = dict(width=w, height=h, samples_per_pixel=256, path_depth=16, use_nee=1, rr_start_depth=1, seed=4)
    hdr, _, _ = o.render(st, s.view)
    cf.check_corridor(hdr, s.view, w, h, 16, 256, 1)


@pytest.mark.parametrize("material", [GGX_MATERIALS[0], GGX_MATERIALS[1], GGX_MATERIALS[2]])
@pytest.mark.parametrize("nee", [1, 0])
def test_ggx_floor_matches_quadrature(oracle_lib, material, nee):
    """GGX floor under a large emitter: the image mean equals Le times the
    quadrature of the published BSDF over the emitter's directions, with NEE
    (light sampling through the BSDF, MaterialEval.cu:652-718) and without
    (BSDF sampling, :720-836)."""
    s = cf.ggx_floor(material)   # Le 0.1: no light sample reaches the 20-unit shadow clamp
    o = oracle_lib.OracleScene(s)
    w = h = 16
    st = dict(width=w, height=h, samples_per_pixel=1024, path_depth=2, use_nee=nee, rr_start_depth=100, seed=11)
    hdr, _, _ = o.render(st, s.view)
    cf.check_ggx_floor(hdr, s.view, w, h, material)
