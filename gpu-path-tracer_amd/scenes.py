"""Procedural scenes of the benchmark configurations, as flat world-space arrays.

Every generator is deterministic (numpy PCG64 with a fixed seed) and returns a
`Scene` whose arrays map one to one onto `pt_scene_desc` (include/pt_core.h).
The reference loads its scenes from YAML + OBJ assets that it does not ship
(.gitignore:18-22 of the reference); these generators replace that asset path.
"""
from dataclasses import dataclass, field
import math

import numpy as np

MODEL_NORMAL, MODEL_DIFFUSE, MODEL_MIRROR, MODEL_PHONG, MODEL_GGX, MODEL_EMISSIVE = range(6)

MATERIAL_DTYPE = np.dtype([
    ("model", "<u4"), ("color", "<f4", 3), ("specular", "<f4", 3), ("roughness", "<f4"),
    ("metallic", "<f4"), ("ior", "<f4"), ("emission", "<f4", 3), ("normal_ts", "<f4", 3),
])
assert MATERIAL_DTYPE.itemsize == 64


def material(model, color=(1.0, 1.0, 1.0), specular=(0.0, 0.0, 0.0), roughness=1.0, metallic=0.0, ior=1.0,
             emission=(0.0, 0.0, 0.0), normal_ts=(0.0, 0.0, 1.0)):
    """One pt_material record.  Defaults follow MaterialDefaults of the reference
    (Core/Include/Core/Graphics/Common/Material.hpp:17-37): white, specular 0,
    rma (1, 0, 1), ior 1; normal_ts (0,0,1) is a flat normal map."""
    m = np.zeros((), MATERIAL_DTYPE)
    m["model"] = model
    m["color"] = color
    m["specular"] = specular
    m["roughness"] = roughness
    m["metallic"] = metallic
    m["ior"] = ior
    m["emission"] = emission
    m["normal_ts"] = normal_ts
    return m


@dataclass
class View:
    fov_y: float                 # radians
    position: tuple
    rotation: tuple              # quaternion x, y, z, w


@dataclass
class Scene:
    name: str
    positions: np.ndarray        # (T, 3, 3) float32
    normals: np.ndarray          # (T, 3, 3) float32
    tangents: np.ndarray         # (T, 3, 4) float32
    material_ids: np.ndarray     # (T,) uint32
    materials: np.ndarray        # (M,) MATERIAL_DTYPE
    light_weights: np.ndarray = None
    view: View = None
    settings: dict = field(default_factory=dict)

    @property
    def triangle_count(self):
        return int(self.positions.shape[0])

    def bounds(self):
        p = self.positions.reshape(-1, 3)
        return p.min(axis=0), p.max(axis=0)


def look_at_quaternion(position, target, up=(0.0, 1.0, 0.0)):
    """Unit quaternion (x, y, z, w) whose rotation maps the camera's local
    forward (0,0,-1), right (1,0,0) and up (0,1,0) onto a look-at frame."""
    p = np.asarray(position, np.float64)
    f = np.asarray(target, np.float64) - p
    f /= np.linalg.norm(f)
    r = np.cross(f, np.asarray(up, np.float64))
    r /= np.linalg.norm(r)
    u = np.cross(r, f)
    m = np.stack([r, u, -f], axis=1)   # columns: images of x, y, z
    tr = m[0, 0] + m[1, 1] + m[2, 2]
    if tr > 0:
        s = math.sqrt(tr + 1.0) * 2
        w, x, y, z = 0.25 * s, (m[2, 1] - m[1, 2]) / s, (m[0, 2] - m[2, 0]) / s, (m[1, 0] - m[0, 1]) / s
    elif m[0, 0] > m[1, 1] and m[0, 0] > m[2, 2]:
        s = math.sqrt(1.0 + m[0, 0] - m[1, 1] - m[2, 2]) * 2
        w, x, y, z = (m[2, 1] - m[1, 2]) / s, 0.25 * s, (m[0, 1] + m[1, 0]) / s, (m[0, 2] + m[2, 0]) / s
    elif m[1, 1] > m[2, 2]:
        s = math.sqrt(1.0 + m[1, 1] - m[0, 0] - m[2, 2]) * 2
        w, x, y, z = (m[0, 2] - m[2, 0]) / s, (m[0, 1] + m[1, 0]) / s, 0.25 * s, (m[1, 2] + m[2, 1]) / s
    else:
        s = math.sqrt(1.0 + m[2, 2] - m[0, 0] - m[1, 1]) * 2
        w, x, y, z = (m[1, 0] - m[0, 1]) / s, (m[0, 2] + m[2, 0]) / s, (m[1, 2] + m[2, 1]) / s, 0.25 * s
    q = np.array([x, y, z, w])
    q /= np.linalg.norm(q)
    return tuple(float(c) for c in q)


def _quad(p0, p1, p2, p3, normal):
    """Two triangles (p0,p1,p2), (p0,p2,p3) with a flat normal and edge tangent."""
    P = np.array([[p0, p1, p2], [p0, p2, p3]], np.float32)
    n = np.asarray(normal, np.float32)
    N = np.broadcast_to(n, (2, 3, 3)).copy()
    t = np.asarray(p1, np.float64) - np.asarray(p0, np.float64)
    t /= np.linalg.norm(t)
    T = np.zeros((2, 3, 4), np.float32)
    T[..., :3] = t
    T[..., 3] = 1.0
    return P, N, T


def _grid_quad(p0, p1, p2, p3, normal, k):
    """The quad (p0, p1, p2, p3) of _quad cut into k x k cells, 2 k^2 triangles
    with the same orientation (cell (i, j) spans p0 + (i, j)/k of the edges
    p1 - p0 and p3 - p0; the quads here are parallelograms)."""
    if k == 1:
        return _quad(p0, p1, p2, p3, normal)
    a = np.asarray(p0, np.float64)
    e1 = np.asarray(p1, np.float64) - a
    e2 = np.asarray(p3, np.float64) - a
    i, j = np.meshgrid(np.arange(k), np.arange(k), indexing="ij")
    i, j = i.ravel()[:, None], j.ravel()[:, None]
    c00 = a + e1 * (i / k) + e2 * (j / k)
    c10 = a + e1 * ((i + 1) / k) + e2 * (j / k)
    c11 = a + e1 * ((i + 1) / k) + e2 * ((j + 1) / k)
    c01 = a + e1 * (i / k) + e2 * ((j + 1) / k)
    P = np.empty((2 * k * k, 3, 3), np.float32)
    P[0::2, 0], P[0::2, 1], P[0::2, 2] = c00, c10, c11
    P[1::2, 0], P[1::2, 1], P[1::2, 2] = c00, c11, c01
    n = np.asarray(normal, np.float32)
    N = np.broadcast_to(n, P.shape).copy()
    t = e1 / np.linalg.norm(e1)
    T = np.zeros((P.shape[0], 3, 4), np.float32)
    T[..., :3] = t
    T[..., 3] = 1.0
    return P, N, T


def _box(lo, hi, k=1):
    """12 k^2 triangles of an axis-aligned box, normals pointing out."""
    x0, y0, z0 = lo
    x1, y1, z1 = hi
    faces = [
        ((x0, y0, z1), (x1, y0, z1), (x1, y1, z1), (x0, y1, z1), (0, 0, 1)),
        ((x1, y0, z0), (x0, y0, z0), (x0, y1, z0), (x1, y1, z0), (0, 0, -1)),
        ((x0, y0, z0), (x0, y0, z1), (x0, y1, z1), (x0, y1, z0), (-1, 0, 0)),
        ((x1, y0, z1), (x1, y0, z0), (x1, y1, z0), (x1, y1, z1), (1, 0, 0)),
        ((x0, y1, z1), (x1, y1, z1), (x1, y1, z0), (x0, y1, z0), (0, 1, 0)),
        ((x0, y0, z0), (x1, y0, z0), (x1, y0, z1), (x0, y0, z1), (0, -1, 0)),
    ]
    parts = [_grid_quad(*f, k) for f in faces]
    return [np.concatenate([p[i] for p in parts]) for i in range(3)]


def cornell_box(tessellation=1):
    """Config 1: box room of 10 wall triangles (floor, ceiling, back, left, right;
    open toward the camera), two axis-aligned boxes (24 triangles) and an
    emissive ceiling quad (2 triangles, radiance 15).  36 triangles.
    tessellation k cuts every quad into k x k cells: the same surfaces in
    36 k^2 triangles (the config-1 generator scaled)."""
    k = int(tessellation)
    white, red, green = 0, 1, 2
    mats = np.stack([
        material(MODEL_DIFFUSE, color=(0.73, 0.73, 0.73)),
        material(MODEL_DIFFUSE, color=(0.65, 0.05, 0.05)),
        material(MODEL_DIFFUSE, color=(0.12, 0.45, 0.15)),
        material(MODEL_EMISSIVE, color=(1.0, 1.0, 1.0), emission=(15.0, 15.0, 15.0)),
    ])
    P, N, T, M = [], [], [], []

    def add(parts, mid):
        P.append(parts[0]); N.append(parts[1]); T.append(parts[2])
        M.append(np.full(parts[0].shape[0], mid, np.uint32))

    add(_grid_quad((-1, 0, -1), (-1, 0, 1), (1, 0, 1), (1, 0, -1), (0, 1, 0), k), white)      # floor
    add(_grid_quad((-1, 2, -1), (1, 2, -1), (1, 2, 1), (-1, 2, 1), (0, -1, 0), k), white)     # ceiling
    add(_grid_quad((-1, 0, -1), (1, 0, -1), (1, 2, -1), (-1, 2, -1), (0, 0, 1), k), white)    # back
    add(_grid_quad((-1, 0, 1), (-1, 0, -1), (-1, 2, -1), (-1, 2, 1), (1, 0, 0), k), red)      # left
    add(_grid_quad((1, 0, -1), (1, 0, 1), (1, 2, 1), (1, 2, -1), (-1, 0, 0), k), green)      # right
    add(_box((0.1, 0.0, -0.1), (0.7, 0.6, 0.5), k), white)                                   # short box
    add(_box((-0.7, 0.0, -0.7), (-0.1, 1.2, -0.1), k), white)                                # tall box
    add(_grid_quad((-0.25, 1.999, -0.25), (0.25, 1.999, -0.25), (0.25, 1.999, 0.25), (-0.25, 1.999, 0.25),
                   (0, -1, 0), k), 3)                                                          # light
    scene = Scene(
        name=f"cornell_box_{36 * k * k}",
        positions=np.concatenate(P), normals=np.concatenate(N), tangents=np.concatenate(T),
        material_ids=np.concatenate(M), materials=mats,
        view=View(fov_y=math.radians(40.0), position=(0.0, 1.0, 3.8), rotation=(0.0, 0.0, 0.0, 1.0)),
        settings=dict(width=256, height=256, samples_per_pixel=16, path_depth=8, use_nee=1,
                      rr_start_depth=3, seed=1),
    )
    assert scene.triangle_count == 36 * k * k
    return scene


def specular_room():
    """The config-1 room with GGX and mirror surfaces where the light lands:
    a copper-like GGX floor, GGX side walls (a red rough dielectric, a green
    half-metal), a mirror back wall, a smooth GGX metal short box and a tinted
    mirror tall box; only the ceiling stays diffuse.  36 triangles."""
    s = cornell_box()
    mats = np.stack([
        material(MODEL_DIFFUSE, color=(0.73, 0.73, 0.73)),                                       # 0 ceiling
        material(MODEL_GGX, color=(0.65, 0.05, 0.05), roughness=0.5, metallic=0.0, ior=1.5),      # 1 left
        material(MODEL_GGX, color=(0.12, 0.45, 0.15), roughness=0.7, metallic=0.3, ior=1.5),     # 2 right
        material(MODEL_EMISSIVE, color=(1.0, 1.0, 1.0), emission=(15.0, 15.0, 15.0)),            # 3 light
        material(MODEL_GGX, color=(0.95, 0.64, 0.54), roughness=0.25, metallic=1.0, ior=1.5),    # 4 floor
        material(MODEL_MIRROR, specular=(0.95, 0.95, 0.95)),                                      # 5 back wall
        material(MODEL_GGX, color=(0.9, 0.9, 0.9), roughness=0.08, metallic=1.0, ior=1.5),       # 6 short box
        material(MODEL_MIRROR, specular=(0.8, 0.9, 1.0)),                                         # 7 tall box
    ])
    mid = s.material_ids.copy()
    mid[0:2] = 4          # floor
    mid[4:6] = 5          # back wall
    mid[10:22] = 6        # short box
    mid[22:34] = 7        # tall box
    s.material_ids = mid
    s.materials = mats
    s.name = "specular_room_36"
    return s


def cornell_scaled(triangles):
    """The config-1 room tessellated to about `triangles` triangles."""
    return cornell_box(max(1, int(round(math.sqrt(triangles / 36.0)))))


_ICO_CACHE = {}


def icosphere(subdivisions):
    """Unit icosphere as a triangle soup (20 * 4^s, 3, 3).  Children of a face
    stay contiguous, so base face f owns rows [f*4^s, (f+1)*4^s)."""
    if subdivisions in _ICO_CACHE:
        return _ICO_CACHE[subdivisions]
    t = (1.0 + math.sqrt(5.0)) / 2.0
    v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t],
                  [0, -1, -t], [0, 1, -t], [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], np.float64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4],
                  [11, 10, 2], [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8],
                  [3, 8, 9], [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]])
    tri = v[f]                                           # (20, 3, 3)
    for _ in range(subdivisions):
        a, b, c = tri[:, 0], tri[:, 1], tri[:, 2]
        ab, bc, ca = a + b, b + c, c + a
        ab /= np.linalg.norm(ab, axis=1, keepdims=True)
        bc /= np.linalg.norm(bc, axis=1, keepdims=True)
        ca /= np.linalg.norm(ca, axis=1, keepdims=True)
        kids = np.stack([np.stack([a, ab, ca], 1), np.stack([ab, b, bc], 1),
                         np.stack([ca, bc, c], 1), np.stack([ab, bc, ca], 1)], 1)
        tri = kids.reshape(-1, 3, 3)
    _ICO_CACHE[subdivisions] = tri
    return tri


def _sphere_tangents(n):
    up = np.zeros_like(n)
    up[..., 1] = 1.0
    alt = np.zeros_like(n)
    alt[..., 0] = 1.0
    axis = np.where(np.abs(n[..., 1:2]) < 0.999, up, alt)
    t = np.cross(axis, n)
    t /= np.linalg.norm(t, axis=-1, keepdims=True)
    return t


def spheres_scene(count=12, subdivisions=6, lights=1024, seed=2, extent=100.0, radius=(2.0, 8.0),
                  light_power_scale=40.0, name=None):
    """Config 2: `count` icospheres (fp32 vertices, uniform-random centres in an
    extent^3 box, radii U[2, 8]) over a 2-triangle ground plane.  Materials per
    base face of each sphere: 40 % diffuse (albedo U[0.1, 0.9]), 40 % GGX
    (roughness U[0.05, 0.8], F0 U[0.04, 1.0] as ior 1.5 + metallic), 20 %
    mirror.  `lights` random sphere triangles are emissive with log-normal(0,1)
    power; the alias table weights them by area, as the reference does."""
    rng = np.random.default_rng(seed)
    unit = icosphere(subdivisions)
    per = unit.shape[0]
    per_face = per // 20
    centers = rng.uniform(0.0, extent, size=(count, 3))
    radii = rng.uniform(radius[0], radius[1], size=count)
    P = (unit[None] * radii[:, None, None, None] + centers[:, None, None, :]).reshape(-1, 3, 3)
    N = np.broadcast_to(unit[None], (count,) + unit.shape).reshape(-1, 3, 3)
    Tn = _sphere_tangents(N)
    T = np.concatenate([Tn, np.ones(Tn.shape[:-1] + (1,))], axis=-1)

    mats = []
    patch_mat = np.empty(count * 20, np.uint32)
    kind = rng.choice(3, size=count * 20, p=[0.4, 0.4, 0.2])
    for i, k in enumerate(kind):
        if k == 0:
            mats.append(material(MODEL_DIFFUSE, color=tuple(rng.uniform(0.1, 0.9, 3))))
        elif k == 1:
            f0 = rng.uniform(0.04, 1.0)
            mats.append(material(MODEL_GGX, color=(1.0, 1.0, 1.0), roughness=float(rng.uniform(0.05, 0.8)),
                                 metallic=float((f0 - 0.04) / 0.96), ior=1.5))
        else:
            mats.append(material(MODEL_MIRROR, specular=(1.0, 1.0, 1.0)))
        patch_mat[i] = i
    mid = np.repeat(patch_mat, per_face)

    # ground plane below the spheres
    y0 = float(centers[:, 1].min() - radii.max() - 2.0)
    g = 10.0 * extent
    gp = _quad((-g, y0, -g), (-g, y0, g + extent), (g + extent, y0, g + extent), (g + extent, y0, -g), (0, 1, 0))
    ground_mat = len(mats)
    mats.append(material(MODEL_DIFFUSE, color=(0.5, 0.5, 0.5)))

    # emissive triangles with log-normal power
    nl = min(lights, P.shape[0])
    light_tris = rng.choice(P.shape[0], size=nl, replace=False)
    power = rng.lognormal(0.0, 1.0, size=nl) * light_power_scale
    e1 = P[light_tris, 1] - P[light_tris, 0]
    e2 = P[light_tris, 2] - P[light_tris, 0]
    area = 0.5 * np.linalg.norm(np.cross(e1, e2), axis=1)
    radiance = power / (math.pi * area)
    for j, t in enumerate(light_tris):
        mid[t] = len(mats)
        mats.append(material(MODEL_EMISSIVE, emission=(float(radiance[j]),) * 3))

    positions = np.concatenate([P, gp[0]]).astype(np.float32)
    normals = np.concatenate([N, gp[1]]).astype(np.float32)
    tangents = np.concatenate([T, gp[2]]).astype(np.float32)
    material_ids = np.concatenate([mid, np.full(2, ground_mat, np.uint32)]).astype(np.uint32)
    cam_pos = (extent * 0.5, extent * 0.6, -1.2 * extent)
    target = (extent * 0.5, extent * 0.4, extent * 0.5)
    return Scene(
        name=name or f"spheres_{count}x{per}",
        positions=positions, normals=normals, tangents=tangents, material_ids=material_ids,
        materials=np.stack(mats),
        view=View(fov_y=math.radians(45.0), position=cam_pos, rotation=look_at_quaternion(cam_pos, target)),
        settings=dict(width=1920, height=1080, samples_per_pixel=64, path_depth=8, use_nee=1,
                      rr_start_depth=3, seed=1),
    )


def production_scene():
    """Config 2 at full size: 12 icospheres at subdivision 6 + ground = 983,042 triangles."""
    s = spheres_scene(count=12, subdivisions=6, lights=1024, seed=2, name="spheres_983042")
    assert s.triangle_count == 983042
    return s


def random_rays(n, lo, hi, seed, t_min=1e-4, t_max=1e34):
    """Oracle ray batch: origins uniform in the box [lo, hi], directions uniform
    on the sphere (z = 1 - 2u0, phi = 2 pi u1).  Returns (n, 8) float32 rows
    (origin, t_min, direction, t_max), the pt_ray layout."""
    rng = np.random.default_rng(seed)
    o = rng.uniform(lo, hi, size=(n, 3))
    u0, u1 = rng.random(n), rng.random(n)
    z = 1.0 - 2.0 * u0
    r = np.sqrt(np.maximum(0.0, 1.0 - z * z))
    phi = 2.0 * np.pi * u1
    d = np.stack([r * np.cos(phi), r * np.sin(phi), z], 1)
    rays = np.empty((n, 8), np.float32)
    rays[:, 0:3] = o
    rays[:, 3] = t_min
    rays[:, 4:7] = d
    rays[:, 7] = t_max
    return rays


def shadow_rays(scene, n, seed):
    """Segments from random surface points to random points on emissive
    triangles, shaped like the NEE shadow rays (t_max = distance - 1e-4)."""
    rng = np.random.default_rng(seed)
    T = scene.triangle_count
    emissive = np.nonzero(scene.materials["model"][scene.material_ids] == MODEL_EMISSIVE)[0]
    src = rng.integers(0, T, n)
    dst = emissive[rng.integers(0, len(emissive), n)]

    def point(tris):
        a, b = rng.random(len(tris)), rng.random(len(tris))
        s = np.sqrt(a)
        b1, b2 = s * (1 - b), s * b
        p = scene.positions[tris].astype(np.float64)
        return p[:, 0] + b1[:, None] * (p[:, 1] - p[:, 0]) + b2[:, None] * (p[:, 2] - p[:, 0])

    o, q = point(src), point(dst)
    v = q - o
    dist = np.linalg.norm(v, axis=1)
    keep = dist > 1e-3
    o, v, dist = o[keep], v[keep], dist[keep]
    rays = np.empty((len(o), 8), np.float32)
    rays[:, 0:3] = o
    rays[:, 3] = 1e-4
    rays[:, 4:7] = v / dist[:, None]
    rays[:, 7] = dist - 1e-4
    return rays


# ---------------------------------------------------------------- batching workload

_C1_COLORS = ((0.73, 0.73, 0.73), (0.65, 0.05, 0.05), (0.12, 0.45, 0.15))


def _flat_frames(P):
    """Per-corner geometric normals and edge tangents of a triangle soup."""
    e1 = (P[:, 1] - P[:, 0]).astype(np.float64)
    e2 = (P[:, 2] - P[:, 0]).astype(np.float64)
    n = np.cross(e1, e2)
    n /= np.maximum(np.linalg.norm(n, axis=1, keepdims=True), 1e-30)
    t = e1 / np.maximum(np.linalg.norm(e1, axis=1, keepdims=True), 1e-30)
    N = np.repeat(n[:, None, :], 3, axis=1).astype(np.float32)
    T = np.ones(P.shape[:2] + (4,), np.float32)
    T[..., :3] = np.repeat(t[:, None, :], 3, axis=1)
    return N, T


def batch_scene(kind, triangles, seed, extent=100.0):
    """One scene of the batching workload: a random triangle soup ("soup") or a
    cluster of tessellated spheres ("spheres") of about `triangles` triangles in
    an extent^3 box, with the config-1 material mix (Lambertian white, red and
    green in 3:1:1) and one emissive quad of radiance 15 above the box."""
    rng = np.random.default_rng(seed)
    if kind == "soup":
        n = int(triangles)
        c = rng.uniform(0.0, extent, size=(n, 1, 3))
        size = 2.5 * extent / n ** (1.0 / 3.0)
        P = (c + rng.normal(0.0, size, size=(n, 3, 3))).astype(np.float32)
        N, T = _flat_frames(P)
        groups = np.arange(n) // 64
    elif kind == "spheres":
        sub = 2
        while 20 * 4 ** (sub + 1) * 4 <= triangles and sub < 7:
            sub += 1
        unit = icosphere(sub)
        per = unit.shape[0]
        count = max(1, int(round(triangles / per)))
        centers = rng.uniform(0.0, extent, size=(count, 3))
        radii = rng.uniform(0.02, 0.08, size=count) * extent
        P = (unit[None] * radii[:, None, None, None] + centers[:, None, None, :]).reshape(-1, 3, 3).astype(np.float32)
        N = np.broadcast_to(unit[None], (count,) + unit.shape).reshape(-1, 3, 3).astype(np.float32)
        Tn = _sphere_tangents(N)
        T = np.concatenate([Tn, np.ones(Tn.shape[:-1] + (1,))], axis=-1).astype(np.float32)
        groups = np.arange(P.shape[0]) // per
    else:
        raise ValueError(kind)
    mats = [material(MODEL_DIFFUSE, color=col) for col in _C1_COLORS]
    mats.append(material(MODEL_EMISSIVE, color=(1.0, 1.0, 1.0), emission=(15.0, 15.0, 15.0)))
    pick = rng.choice(3, size=int(groups.max()) + 1, p=[0.6, 0.2, 0.2]).astype(np.uint32)
    mid = pick[groups]
    top = float(P[..., 1].max()) + 0.1 * extent
    lo, hi = 0.2 * extent, 0.8 * extent
    lq = _quad((lo, top, lo), (hi, top, lo), (hi, top, hi), (lo, top, hi), (0, -1, 0))
    positions = np.concatenate([P, lq[0]]).astype(np.float32)
    normals = np.concatenate([N, lq[1]]).astype(np.float32)
    tangents = np.concatenate([T, lq[2]]).astype(np.float32)
    material_ids = np.concatenate([mid, np.full(2, 3, np.uint32)]).astype(np.uint32)
    cam = (0.5 * extent, 0.6 * extent, -0.9 * extent)
    target = (0.5 * extent, 0.4 * extent, 0.5 * extent)
    return Scene(
        name=f"{kind}_{positions.shape[0]}", positions=positions, normals=normals, tangents=tangents,
        material_ids=material_ids, materials=np.stack(mats),
        view=View(fov_y=math.radians(45.0), position=cam, rotation=look_at_quaternion(cam, target)),
        settings=dict(width=512, height=512, samples_per_pixel=16, path_depth=8, use_nee=1, rr_start_depth=3, seed=1),
    )


BATCH_RESOLUTIONS = ((256, 256), (512, 512), (640, 480), (1280, 720), (1920, 1080))
BATCH_SPP = (4, 9, 16, 36, 64)


def batch_workload(count=32, seed=3, tri_range=(10_000, 2_000_000), extent=100.0):
    """Config 3: `count` scenes with log-uniform triangle counts in tri_range,
    soups and sphere clusters alternating, each seen from 1-4 cameras at
    resolutions 256^2..1920x1080 and 4..64 spp.  Returns (scenes, views) with
    views = [(scene index, settings, View)]."""
    rng = np.random.default_rng(seed)
    tris = np.exp(rng.uniform(math.log(tri_range[0]), math.log(tri_range[1]), count)).astype(np.int64)
    out_scenes, views = [], []
    for i, t in enumerate(tris):
        kind = "soup" if i % 2 == 0 else "spheres"
        out_scenes.append(batch_scene(kind, int(t), seed * 1000 + i, extent))
        for _ in range(int(rng.integers(1, 5))):
            w, h = BATCH_RESOLUTIONS[int(rng.integers(len(BATCH_RESOLUTIONS)))]
            spp = BATCH_SPP[int(rng.integers(len(BATCH_SPP)))]
            ang = rng.uniform(0.0, 2.0 * math.pi)
            pos = (0.5 * extent + 1.4 * extent * math.cos(ang), rng.uniform(0.3, 0.9) * extent,
                   0.5 * extent + 1.4 * extent * math.sin(ang))
            target = (0.5 * extent, 0.4 * extent, 0.5 * extent)
            v = View(fov_y=math.radians(45.0), position=pos, rotation=look_at_quaternion(pos, target))
            st = dict(width=w, height=h, samples_per_pixel=spp, path_depth=8, use_nee=1, rr_start_depth=3,
                      seed=int(rng.integers(1, 1 << 31)))
            views.append((i, st, v))
    return out_scenes, views


# ---------------------------------------------------------------- serving workload

SERVE_RESOLUTIONS = ((512, 512), (1280, 720), (1920, 1080), (2560, 1440), (3840, 2160))
SERVE_SPP = (4, 9, 16, 36, 64, 121, 256)


def serve_request(rng, scene_count, resolutions=SERVE_RESOLUTIONS, spp=SERVE_SPP, depth=8):
    """One request of config 4: a random resident scene, a camera inside the
    config-1 room, a resolution and an spp (square counts: the reference rounds
    spp down to a square)."""
    sid = int(rng.integers(scene_count))
    w, h = resolutions[int(rng.integers(len(resolutions)))]
    s = spp[int(rng.integers(len(spp)))]
    pos = (rng.uniform(-0.6, 0.6), rng.uniform(0.6, 1.6), rng.uniform(1.5, 3.8))
    target = (rng.uniform(-0.5, 0.5), rng.uniform(0.3, 1.2), rng.uniform(-1.0, 0.0))
    v = View(fov_y=math.radians(40.0), position=pos, rotation=look_at_quaternion(pos, target))
    st = dict(width=w, height=h, samples_per_pixel=s, path_depth=depth, use_nee=1, rr_start_depth=3,
              seed=int(rng.integers(1, 1 << 31)))
    return sid, st, v


# ---------------------------------------------------------------- guiding workload

def guiding_room(triangles=200_000, seed=5):
    """Config 5: two rooms split by a 10 cm wall at z = 0 with a 0.5 m gap
    (x in [1.0, 1.5], full height).  The only light is 4 emissive triangles
    (two 0.4 m quads, radiance 50) under the ceiling of the back room, out of
    sight of the camera in the front room.  Walls Lambertian (config-1 white,
    red, green), two GGX boxes in the front room; every quad but the light is
    tessellated so the scene holds about `triangles` triangles."""
    quads = []   # (p0, p1, p2, p3, normal, material)
    white, red, green, ggx, light = 0, 1, 2, 3, 4
    X0, X1, Y1, Z0, Z1 = -2.0, 2.0, 2.5, -4.0, 3.0
    # outer box, normals inward
    quads += [((X0, 0, Z0), (X0, 0, Z1), (X1, 0, Z1), (X1, 0, Z0), (0, 1, 0), white),        # floor
              ((X0, Y1, Z0), (X1, Y1, Z0), (X1, Y1, Z1), (X0, Y1, Z1), (0, -1, 0), white),    # ceiling
              ((X0, 0, Z0), (X1, 0, Z0), (X1, Y1, Z0), (X0, Y1, Z0), (0, 0, 1), white),       # back
              ((X1, 0, Z1), (X0, 0, Z1), (X0, Y1, Z1), (X1, Y1, Z1), (0, 0, -1), white),      # front
              ((X0, 0, Z1), (X0, 0, Z0), (X0, Y1, Z0), (X0, Y1, Z1), (1, 0, 0), red),         # left
              ((X1, 0, Z0), (X1, 0, Z1), (X1, Y1, Z1), (X1, Y1, Z0), (-1, 0, 0), green)]      # right

    def slab(xa, xb, za, zb):
        """Faces of the wall piece [xa, xb] x [0, Y1] x [za, zb] facing the rooms and the gap."""
        return [((xa, 0, zb), (xb, 0, zb), (xb, Y1, zb), (xa, Y1, zb), (0, 0, 1), white),
                ((xb, 0, za), (xa, 0, za), (xa, Y1, za), (xb, Y1, za), (0, 0, -1), white),
                ((xa, 0, za), (xa, 0, zb), (xa, Y1, zb), (xa, Y1, za), (-1, 0, 0), white),
                ((xb, 0, zb), (xb, 0, za), (xb, Y1, za), (xb, Y1, zb), (1, 0, 0), white)]

    quads += slab(X0, 1.0, -0.05, 0.05)[:2] + [slab(X0, 1.0, -0.05, 0.05)[3]]
    quads += slab(1.5, X1, -0.05, 0.05)[:3]
    s = 0.4
    lights = [((-1.4, Y1 - 0.01, -3.3), (-1.4 + s, Y1 - 0.01, -3.3), (-1.4 + s, Y1 - 0.01, -3.3 + s),
               (-1.4, Y1 - 0.01, -3.3 + s), (0, -1, 0), light),
              ((-0.8, Y1 - 0.01, -3.3), (-0.8 + s, Y1 - 0.01, -3.3), (-0.8 + s, Y1 - 0.01, -3.3 + s),
               (-0.8, Y1 - 0.01, -3.3 + s), (0, -1, 0), light)]
    boxes = [((-1.2, 0.0, 0.8), (-0.4, 0.9, 1.6)), ((0.3, 0.0, 0.4), (0.9, 1.4, 1.0))]
    box_faces = 6 * len(boxes)
    k = max(1, int(round(math.sqrt((triangles - 4) / (2.0 * (len(quads) + box_faces))))))
    mats = np.stack([
        material(MODEL_DIFFUSE, color=_C1_COLORS[0]), material(MODEL_DIFFUSE, color=_C1_COLORS[1]),
        material(MODEL_DIFFUSE, color=_C1_COLORS[2]),
        material(MODEL_GGX, color=(0.9, 0.8, 0.6), roughness=0.3, metallic=0.0, ior=1.5),
        material(MODEL_EMISSIVE, color=(1.0, 1.0, 1.0), emission=(50.0, 50.0, 50.0)),
    ])
    P, N, T, M = [], [], [], []
    for *pts, nrm, mid in quads:
        g = _grid_quad(*pts, nrm, k)
        P.append(g[0]); N.append(g[1]); T.append(g[2]); M.append(np.full(len(g[0]), mid, np.uint32))
    for lo, hi in boxes:
        b = _box(lo, hi, k)
        P.append(b[0]); N.append(b[1]); T.append(b[2]); M.append(np.full(len(b[0]), ggx, np.uint32))
    for *pts, nrm, mid in lights:
        g = _quad(*pts, nrm)
        P.append(g[0]); N.append(g[1]); T.append(g[2]); M.append(np.full(2, light, np.uint32))
    cam = (-0.2, 1.4, 2.8)
    target = (0.6, 1.0, 0.0)
    return Scene(
        name=f"guiding_room_{sum(len(p) for p in P)}",
        positions=np.concatenate(P).astype(np.float32), normals=np.concatenate(N).astype(np.float32),
        tangents=np.concatenate(T).astype(np.float32), material_ids=np.concatenate(M), materials=mats,
        view=View(fov_y=math.radians(55.0), position=cam, rotation=look_at_quaternion(cam, target)),
        settings=dict(width=1280, height=720, samples_per_pixel=64, path_depth=8, use_nee=1, rr_start_depth=3,
                      seed=1),
    )
