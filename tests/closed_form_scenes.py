"""Scenes whose rendered radiance has a closed form, and those closed forms.

Shared by the CPU tests of the oracle (tests/test_shading_closed_forms.py)
and the GPU tests of the core (tests/test_gpu_shading.py).

The reference's material model puts emission only on Emissive triangles, which
end a path (MaterialEval.cu:882-912), so the textbook furnace (walls that both
emit and reflect) cannot be built with it.  What can, and what these scenes
use:

* plates: a diffuse floor of albedo rho under a parallel emissive rectangle of
  radiance Le.  Every path is camera -> floor -> emitter or escape, so the
  expected radiance of a floor point p is rho * Le * F(p), F the form factor
  from a differential area at p to the rectangle (closed form, Howell's
  catalogue, configuration B-3), with NEE on or off and with or without
  Russian roulette;
* corridor: two parallel mirror walls of reflectance s (per channel) and an
  emitter across the far end.  A camera ray reflects k times, k fixed by the
  unfolded geometry, so its radiance is exactly s^k Le when k < depth limit
  (mirrors are delta: the emitter's radiance counts with NEE on as well);
* ggx floor: a GGX floor under a large emissive rectangle; the expected
  radiance is Le * integral of f(wo, wi) cos(theta_i) over the directions that
  reach the rectangle, computed by quadrature from the published GGX formulas
  (Walter et al. 2007; Heitz 2014 for the Smith term) in float64;
* sphere furnace: the inside of a sphere whose triangles are diffuse (albedo
  rho) or emissive (radiance Le, a fraction f of the area).  Between two points
  x, y of a sphere of radius R, cos(theta_x) = cos(theta_y) = |x - y| / (2R),
  so the form factor from x to an area dA is dA / (4 pi R^2) wherever x and dA
  lie: a cosine-sampled bounce lands on emissive area with probability f, and a
  light sample by area carries rho f Le exactly.  A diffuse point then reads the
  multi-bounce series rho f Le (1 + q + ... + q^(n-1)), q = rho (1 - f), with
  n = depth bounces' light samples under NEE and n = depth - 1 emitter hits
  without it: the diffuse furnace of the material model.
"""
import math

import numpy as np

import scenes


def _scene(name, parts, mats, view, settings):
    P = np.concatenate([p[0] for p, _ in parts]).astype(np.float32)
    N = np.concatenate([p[1] for p, _ in parts]).astype(np.float32)
    T = np.concatenate([p[2] for p, _ in parts]).astype(np.float32)
    M = np.concatenate([np.full(p[0].shape[0], m, np.uint32) for p, m in parts])
    return scenes.Scene(name=name, positions=P, normals=N, tangents=T, material_ids=M, materials=np.stack(mats),
                        view=view, settings=settings)


# ---------------------------------------------------------------- plates

PLATE_H = 1.0          # emitter height above the floor
PLATE_S = 2.0          # emitter half size: the rectangle [-S, S] x [-S, S] at y = H
FLOOR_S = 50.0         # floor half size (paths only need the floor under the camera)


def plates(rho=0.6, le=3.0, emitter_half=PLATE_S):
    """Diffuse floor (albedo rho) at y = 0 under an emissive square of half
    size `emitter_half` at y = PLATE_H; camera at y = 0.5 looking straight down."""
    S, H, F = emitter_half, PLATE_H, FLOOR_S
    floor = scenes._quad((-F, 0, -F), (-F, 0, F), (F, 0, F), (F, 0, -F), (0, 1, 0))
    light = scenes._quad((-S, H, -S), (S, H, -S), (S, H, S), (-S, H, S), (0, -1, 0))
    mats = [scenes.material(scenes.MODEL_DIFFUSE, color=(rho, rho, rho)),
            scenes.material(scenes.MODEL_EMISSIVE, emission=(le, le, le))]
    pos = (0.0, 0.5, 0.0)
    view = scenes.View(fov_y=math.radians(60.0), position=pos,
                       rotation=scenes.look_at_quaternion(pos, (0.0, 0.0, 0.0), up=(0.0, 0.0, -1.0)))
    return _scene("plates", [(floor, 0), (light, 1)], mats, view, {})


def _corner_ff(a, b, h):
    """Form factor from a differential area to a parallel a x b rectangle at
    height h whose corner lies straight above it (Howell B-3)."""
    A, B = a / h, b / h
    sa, sb = np.sqrt(1.0 + A * A), np.sqrt(1.0 + B * B)
    return (A / sa * np.arctan(B / sa) + B / sb * np.arctan(A / sb)) / (2.0 * math.pi)


def plate_form_factor(x, z, half=PLATE_S, h=PLATE_H):
    """Form factor from floor points (x, 0, z) to the square [-half, half]^2 at
    height h: the four corner rectangles around the foot point, signed."""
    x = np.asarray(x, np.float64)
    z = np.asarray(z, np.float64)
    total = 0.0
    # inclusion-exclusion over the corners [0, a] x [0, b] of signed extents
    for cx, ex in ((half, 1.0), (-half, -1.0)):
        for cz, ez in ((half, 1.0), (-half, -1.0)):
            a, b = cx - x, cz - z
            total = total + ex * ez * np.sign(a) * np.sign(b) * _corner_ff(np.abs(a), np.abs(b), h)
    return total


def floor_points(view, width, height, fx=0.5, fy=0.5):
    """World points where the rays through pixel position (x + fx, y + fy) of
    a camera looking straight down meet the plane y = 0 (numpy float64 camera
    of Renderer.cpp:237-257)."""
    o, llc, hor, ver = camera_basis64(view, width / height)
    x = (np.arange(width) + fx) / width
    y = (np.arange(height) + fy) / height
    u, v = np.meshgrid(x, y)
    d = llc[None, None] + u[..., None] * hor + v[..., None] * ver - o
    t = -o[1] / d[..., 1]
    p = o + t[..., None] * d
    return p.reshape(-1, 3)


def camera_basis64(view, aspect):
    q = np.asarray(view.rotation, np.float64)
    w, qv = q[3], q[:3]

    def rot(v):
        uv = np.cross(qv, v)
        uuv = np.cross(qv, uv)
        return v + 2.0 * (w * uv + uuv)

    f, r, u = rot(np.array([0.0, 0.0, -1.0])), rot(np.array([1.0, 0.0, 0.0])), rot(np.array([0.0, 1.0, 0.0]))
    hh = math.tan(view.fov_y / 2)
    hw = hh * aspect
    o = np.asarray(view.position, np.float64)
    llc = o + f - hw * r - hh * u
    return o, llc, 2 * hw * r, 2 * hh * u


# ---------------------------------------------------------------- corridor

CORRIDOR_HALF = 1.0    # mirror walls at x = -1 and x = +1
CORRIDOR_END = 30.0    # emitter across the corridor at z = -30
CORRIDOR_LE = (6.0, 5.0, 4.0)
CORRIDOR_S = (0.9, 0.6, 0.3)


def corridor(spec=CORRIDOR_S, le=CORRIDOR_LE, yaw=0.12):
    """Mirror walls x = +-1 (y in [-60, 60], z in [-30, 2]) and an emitter across
    z = -30; camera at the origin looking down -z, turned by `yaw` radians."""
    W, E = CORRIDOR_HALF, CORRIDOR_END
    y0, y1, z1 = -60.0, 60.0, 2.0
    left = scenes._quad((-W, y0, z1), (-W, y0, -E), (-W, y1, -E), (-W, y1, z1), (1, 0, 0))
    right = scenes._quad((W, y0, -E), (W, y0, z1), (W, y1, z1), (W, y1, -E), (-1, 0, 0))
    end = scenes._quad((-W, y0, -E), (W, y0, -E), (W, y1, -E), (-W, y1, -E), (0, 0, 1))
    mats = [scenes.material(scenes.MODEL_MIRROR, specular=spec),
            scenes.material(scenes.MODEL_EMISSIVE, emission=le)]
    target = (math.sin(yaw), 0.0, -math.cos(yaw))
    view = scenes.View(fov_y=math.radians(20.0), position=(0.0, 0.0, 0.0),
                       rotation=scenes.look_at_quaternion((0.0, 0.0, 0.0), target))
    return _scene("corridor", [(left, 0), (right, 0), (end, 1)], mats, view, {})


def corridor_bounces(view, width, height, fx=0.5, fy=0.5):
    """Number of wall reflections before the ray through pixel position
    (x + fx, y + fy) reaches the end wall, from the unfolded straight line."""
    o, llc, hor, ver = camera_basis64(view, width / height)
    x = (np.arange(width) + fx) / width
    y = (np.arange(height) + fy) / height
    u, v = np.meshgrid(x, y)
    d = llc[None, None] + u[..., None] * hor + v[..., None] * ver - o
    t = (-CORRIDOR_END - o[2]) / d[..., 2]
    xu = o[0] + t * d[..., 0]          # x where the unfolded line meets the end plane
    k = np.floor((np.abs(xu) + CORRIDOR_HALF) / (2 * CORRIDOR_HALF)).astype(np.int64)
    return k.reshape(-1)


def corridor_expected(view, width, height, depth):
    """(expected rgb per pixel, mask of pixels whose every sample reflects the
    same number of times and stays between the walls)."""
    ks = [corridor_bounces(view, width, height, fx, fy) for fx in (0.0, 1.0) for fy in (0.0, 1.0)]
    same = np.all([k == ks[0] for k in ks], axis=0)
    k = ks[0]
    s = np.asarray(CORRIDOR_S, np.float64)
    le = np.asarray(CORRIDOR_LE, np.float64)
    val = np.where((k < depth)[:, None], le[None] * s[None] ** k[:, None], 0.0)
    return val, same


# ---------------------------------------------------------------- ggx floor

GGX_EMITTER_HALF = 3.0


def ggx_floor(material, le=0.1):
    """GGX floor of `material` = (color, roughness, metallic, ior) at y = 0 under
    an emissive square of half size 3 at y = 1; camera at (0, 0.8, 0.6)
    looking at the floor point (0, 0, 0) (37 degrees from the normal)."""
    color, rough, metal, ior = material
    S, H, F = GGX_EMITTER_HALF, PLATE_H, FLOOR_S
    floor = scenes._quad((-F, 0, -F), (-F, 0, F), (F, 0, F), (F, 0, -F), (0, 1, 0))
    light = scenes._quad((-S, H, -S), (S, H, -S), (S, H, S), (-S, H, S), (0, -1, 0))
    mats = [scenes.material(scenes.MODEL_GGX, color=color, roughness=rough, metallic=metal, ior=ior),
            scenes.material(scenes.MODEL_EMISSIVE, emission=(le, le, le))]
    pos = (0.0, 0.8, 0.6)
    view = scenes.View(fov_y=math.radians(10.0), position=pos, rotation=scenes.look_at_quaternion(pos, (0, 0, 0)))
    return _scene("ggx_floor", [(floor, 0), (light, 1)], mats, view, {})


def ggx_f64(material, wo, wi):
    """BSDF (n, 3) of local directions in float64 from the published forms:
    D = a^2 / (pi cos^4 (a^2 + tan^2)^2), Smith height-correlated
    G = 1 / (1 + L(wo) + L(wi)), L = (sqrt(1 + a^2 tan^2) - 1) / 2, Schlick F."""
    color, rough, metal, ior = material
    a = max(rough * rough, 1e-3)
    base = np.asarray(color, np.float64)
    f0 = ((1.0 - ior) / (1.0 + ior)) ** 2 * (1 - metal) + base * metal
    diff = base * (1.0 - metal)
    wm = wo + wi
    wm /= np.linalg.norm(wm, axis=-1, keepdims=True)
    cm = wm[:, 2]
    t2m = (1 - cm * cm) / (cm * cm)
    D = a * a / (math.pi * cm ** 4 * (a * a + t2m) ** 2)

    def lam(w):
        t2 = (1 - w[:, 2] ** 2) / (w[:, 2] ** 2)
        return (np.sqrt(1 + a * a * t2) - 1) / 2

    G = 1 / (1 + lam(wo) + lam(wi))
    owm = np.clip((wo * wm).sum(-1), 0, 1)
    F = f0[None] + (1 - f0[None]) * ((1 - owm) ** 5)[:, None]
    spec = F * (D * G / (4 * wi[:, 2] * wo[:, 2]))[:, None]
    return (1 - F) * diff[None] / math.pi + spec


def ggx_floor_expected(material, wo, p, le=0.1, n_theta=1200, n_phi=1200):
    """Le * integral over the hemisphere of f(wo, wi) cos(theta_i) for the
    directions from floor point p that reach the emitter, by the midpoint
    rule in (cos^2 theta, phi) (so d(omega) cos = d(c2) d(phi) / 2)."""
    c2 = (np.arange(n_theta) + 0.5) / n_theta
    ph = (np.arange(n_phi) + 0.5) / n_phi * 2 * math.pi
    C2, PH = np.meshgrid(c2, ph, indexing="ij")
    ct = np.sqrt(C2).ravel()
    st = np.sqrt(1 - C2).ravel()
    wi = np.stack([st * np.cos(PH.ravel()), st * np.sin(PH.ravel()), ct], 1)
    # the BSDF is isotropic: any frame with the floor normal (+y) as local z
    # serves; here local (x, y, z) = world (x, z, y), and `wo` is given in it
    world = np.stack([wi[:, 0], wi[:, 2], wi[:, 1]], 1)
    t = (PLATE_H - p[1]) / world[:, 1]
    hx, hz = p[0] + t * world[:, 0], p[2] + t * world[:, 2]
    inside = (np.abs(hx) < GGX_EMITTER_HALF) & (np.abs(hz) < GGX_EMITTER_HALF)
    wol = np.broadcast_to(np.asarray(wo, np.float64), wi.shape)
    f = ggx_f64(material, wol, wi)
    w = (1.0 / n_theta) * (2 * math.pi / n_phi) / 2.0
    return le * (f * inside[:, None]).sum(0) * w


# ---------------------------------------------------------------- sphere furnace

FURNACE_R = 2.0                        # every segment stays below the 20-unit shadow clamp
FURNACE_EMISSIVE = (0, 3, 8, 14, 17)   # base faces of the icosphere that emit: f = 1/4


def sphere_furnace(rho=0.8, le=2.0, subdivisions=5):
    """Inside of an icosphere (20 * 4^s flat triangles, wound and normalled
    inward), the patches of FURNACE_EMISSIVE emissive and the rest diffuse;
    camera at the centre."""
    unit = scenes.icosphere(subdivisions)[:, [0, 2, 1]]
    per = unit.shape[0] // 20
    P = unit * FURNACE_R
    N = -unit
    Tn = scenes._sphere_tangents(N)
    T = np.concatenate([Tn, np.ones(Tn.shape[:-1] + (1,))], axis=-1)
    M = np.zeros(len(P), np.uint32)
    for f in FURNACE_EMISSIVE:
        M[f * per:(f + 1) * per] = 1
    mats = [scenes.material(scenes.MODEL_DIFFUSE, color=(rho, rho, rho)),
            scenes.material(scenes.MODEL_EMISSIVE, emission=(le, le, le))]
    pos = (0.0, 0.0, 0.0)
    view = scenes.View(fov_y=math.radians(70.0), position=pos,
                       rotation=scenes.look_at_quaternion(pos, (0.3, -0.2, -1.0)))
    return scenes.Scene(name="sphere_furnace", positions=P.astype(np.float32), normals=N.astype(np.float32),
                        tangents=T.astype(np.float32), material_ids=M, materials=np.stack(mats), view=view,
                        settings={})


def furnace_fraction(scene):
    """Emissive share f of the furnace's area (float64 over its fp32 triangles)."""
    P = scene.positions.astype(np.float64)
    area = 0.5 * np.linalg.norm(np.cross(P[:, 1] - P[:, 0], P[:, 2] - P[:, 0]), axis=1)
    return float(area[scene.material_ids == 1].sum() / area.sum())


def furnace_expected(rho, le, f, depth, nee):
    """E[L] of a diffuse point: rho f Le sum_{k < n} (rho (1 - f))^k, n = depth
    with NEE, depth - 1 without."""
    n = depth if nee else depth - 1
    q = rho * (1.0 - f)
    return rho * f * le * sum(q ** k for k in range(n))


def furnace_pixels(scene, closest, w, h, material=0):
    """Pixels whose four corner rays all hit triangles of `material` (0
    diffuse, 1 emissive): none straddles an emitter edge.  `closest` maps
    (n, 8) rays to (n, 4) hit records."""
    o, llc, hor, ver = camera_basis64(scene.view, w / h)
    ok = np.ones(w * h, bool)
    for fx in (0.0, 1.0):
        for fy in (0.0, 1.0):
            x, y = np.meshgrid((np.arange(w) + fx) / w, (np.arange(h) + fy) / h)
            d = llc[None, None] + x[..., None] * hor + y[..., None] * ver - o
            d /= np.linalg.norm(d, axis=-1, keepdims=True)
            rays = np.zeros((w * h, 8), np.float32)
            rays[:, 0:3] = o
            rays[:, 3] = 1e-4
            rays[:, 4:7] = d.reshape(-1, 3)
            rays[:, 7] = 1e34
            prim = closest(rays)[:, 0]
            ok &= scene.material_ids[prim] == material
    return ok


def check_furnace(hdr, mask, rho, le, f, depth, nee):
    """The diffuse pixels' mean within 4 standard errors plus 3e-4 relative
    (the flat facets bend the sphere's identity by about 1e-4) of the series;
    with NEE at depth 1 every pixel is within 1e-2 of rho f Le (one light
    sample per sample, carrying rho f Le up to the facets)."""
    want = furnace_expected(rho, le, f, depth, nee)
    g = hdr[mask, :3].astype(np.float64)
    assert np.allclose(g[:, 0], g[:, 1], rtol=1e-5) and np.allclose(g[:, 0], g[:, 2], rtol=1e-5)
    v = g[:, 0]
    se = v.std() / math.sqrt(len(v))
    assert abs(v.mean() - want) < 4 * se + 3e-4 * want, (v.mean(), want, se)
    if nee and depth == 1:
        assert np.allclose(v, want, rtol=1e-2)


# ---------------------------------------------------------------- checks shared by the oracle and GPU tests

def check_plates(hdr, view, w, h, rho, le, nee, depth, blk=8):
    """E[L] = rho Le F(p) per 8 x 8 block within 5 standard errors (0 at depth 1
    without NEE).  hdr: (w h, 4) radiance / spp."""
    p = floor_points(view, w, h)
    want = rho * le * plate_form_factor(p[:, 0], p[:, 2])
    got = hdr[:, 0].astype(np.float64)
    # grey: the channels differ only by the order of their float additions
    assert np.allclose(hdr[:, 0], hdr[:, 1], rtol=1e-5) and np.allclose(hdr[:, 0], hdr[:, 2], rtol=1e-5)
    if depth == 1 and not nee:
        assert (got == 0).all()
        return
    res = (got - want).reshape(h // blk, blk, w // blk, blk)
    mean, se = res.mean((1, 3)), res.std((1, 3)) / blk
    assert (np.abs(mean) < 5 * se + 1e-4).all(), (mean, se)
    assert abs(res.mean()) < 4 * res.std() / math.sqrt(w * h) + 1e-5


def check_corridor(hdr, view, w, h, depth, spp, rr_start):
    """Without roulette every pixel whose samples all reflect k times reads
    s^k Le (0 for k >= depth); with roulette from the first bounce the pixels
    are unbiased around it (z-scores of the exact per-pixel variance)."""
    want, same = corridor_expected(view, w, h, depth)
    ks = corridor_bounces(view, w, h)
    assert same.mean() > 0.6 and len(set(ks[same])) >= 4
    got = hdr[same, :3].astype(np.float64)
    if rr_start >= depth:
        assert np.allclose(got, want[same], rtol=1e-5, atol=1e-6)
        return
    k = ks[same]
    sv = np.asarray(CORRIDOR_S, np.float64)
    le = np.asarray(CORRIDOR_LE, np.float64)
    p = sv.max() ** k
    v = le[None] * (sv / sv.max())[None] ** k[:, None]
    se = np.sqrt(v * v * (p * (1 - p))[:, None] / spp) + 1e-9
    z = (got - want[same]) / se
    assert np.allclose(got[k == 0], want[same][k == 0])
    assert abs(z.mean()) < 0.25 and (np.abs(z) < 6.5).all(), (z.mean(), np.abs(z).max())


def check_ggx_floor(hdr, view, w, h, material, blk=4):
    """Block means against Le times the quadrature of the published BSDF over
    the emitter's directions at each block's centre."""
    img = hdr[:, :3].astype(np.float64).reshape(h // blk, blk, w // blk, blk, 3)
    got, se = img.mean((1, 3)), img.std((1, 3)) / blk
    cam = np.asarray(view.position, np.float64)
    centres = floor_points(view, w // blk, h // blk)
    want = np.zeros_like(got)
    for i, q in enumerate(centres):
        wo = cam - q
        wo /= np.linalg.norm(wo)
        want.reshape(-1, 3)[i] = ggx_floor_expected(material, np.array([wo[0], wo[2], wo[1]]), q, n_theta=800,
                                                    n_phi=800)
    assert (np.abs(got - want) < 5 * se + 3e-3 * want).all(), (got, want, se)
