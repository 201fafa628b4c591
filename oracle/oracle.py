"""ctypes binding of the CPU oracle (test infrastructure only: imported by
tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg)."""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "liboracle.so")
BIN_DTYPE = np.dtype([("q", "<f4"), ("p", "<f4"), ("alias", "<u4"), ("value", "<u4")])

_P = C.c_void_p
_U32 = C.c_uint32
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError(f"{LIB} missing: run `python oracle/build.py`")
        L = C.CDLL(LIB)
        sig = {
            "orc_scene_create": (_P, [_U32, _P, _P, _P, _P, _U32, _P, _P]),
            "orc_scene_destroy": (None, [_P]),
            "orc_scene_info": (None, [_P, _P]),
            "orc_scene_bvh": (None, [_P, _P, _P]),
            "orc_trace_closest": (None, [_P, _P, _U32, _P, C.c_int, _U32, _P, _P]),
            "orc_trace_any": (None, [_P, _P, _U32, _P, C.c_int, _U32]),
            "orc_classify": (None, [_P, _P, _U32, _P, _U32]),
            "orc_bvh4_closest": (None, [_P, _P, _P, _P, _U32, _P, _P, _P, _P, _U32]),
            "orc_bvh4q_any": (None, [_P, _P, _P, _P, _U32, _P, _P, _P, _U32]),
            "orc_bvh8q_any": (None, [_P, _P, _P, _P, _U32, _P, _P, _P, _P, _U32]),
            "orc_bvh8q_closest": (None, [_P, _P, _P, _P, _U32, _P, _P, _P, _P, _U32]),
            "orc_alias_build": (None, [_P, _U32, _P]),
            "orc_alias_sample": (None, [_P, _U32, _P, _U32, _P, _P]),
            "orc_scene_alias": (None, [_P, _P]),
            "orc_pcg32": (None, [C.c_uint64, C.c_uint64, C.c_uint64, _U32, _P]),
            "orc_camera_basis": (None, [C.c_float, C.c_float, _P, _P, _P]),
            "orc_render": (None, [_P, _P, C.c_uint64, _P, _U32, _U32, _U32, _P, _P]),
            "orc_tonemap": (None, [_P, _U32, _P]),
            "orc_trace_path": (_U32, [_P, _P, C.c_uint64, _P, C.c_uint64, _P, _U32, _P]),
            "orc_sincos": (None, [_P, _U32, _P, _P]),
            "orc_cosine_hemisphere": (None, [_P, _U32, _P]),
            "orc_ggx_term": (None, [C.c_int, C.c_float, _P, _U32, _P]),
            "orc_ggx_dvis": (None, [C.c_float, _P, _P, _U32, _P]),
            "orc_ggx_sample_wm": (None, [C.c_float, _P, _P, _U32, _P]),
            "orc_ggx_bsdf": (None, [_P, _P, _P, _U32, _P, _P]),
            "orc_ggx_pdf": (None, [_P, _P, _P, _U32, _P]),
            "orc_ggx_sample": (None, [_P, _P, _P, _U32, _P, _P, _P, _P, _P]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def default_threads():
    return max(1, min(os.cpu_count() or 1, 16))


class OracleScene:
    """CPU restatement of the reference scene upload + traversal + shading."""

    def __init__(self, scene):
        self._arrs = [np.ascontiguousarray(scene.positions, np.float32),
                      np.ascontiguousarray(scene.normals, np.float32),
                      None if scene.tangents is None else np.ascontiguousarray(scene.tangents, np.float32),
                      np.ascontiguousarray(scene.material_ids, np.uint32),
                      np.ascontiguousarray(scene.materials),
                      None if scene.light_weights is None else np.ascontiguousarray(scene.light_weights, np.float32)]
        P, N, T, M, mats, W = self._arrs
        self.h = lib().orc_scene_create(scene.triangle_count, _p(P), _p(N), _p(T), _p(M), len(mats), _p(mats), _p(W))
        info = np.zeros(6, np.uint32)
        lib().orc_scene_info(self.h, _p(info))
        self.info = dict(nodes=int(info[0]), depth=int(info[1]), max_depth=int(info[2]), leaf_size=int(info[3]),
                         lights=int(info[4]), triangles=int(info[5]))

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_scene_destroy(self.h)
            self.h = None

    def closest(self, rays, brute=False, threads=None, counts=False, tprim=False):
        """Closest hits (prim, t, u, v as uint32 words).  Default: the reference's
        traversal order (first found wins at equal t); brute: all triangles under
        the (t, prim id) order; tprim: the same order through the BVH."""
        rays = np.ascontiguousarray(rays, np.float32)
        hits = np.zeros((len(rays), 4), np.uint32)
        nv = np.zeros(len(rays), np.uint32) if counts else None
        nt = np.zeros(len(rays), np.uint32) if counts else None
        mode = 1 if brute else 2 if tprim else 0
        lib().orc_trace_closest(self.h, _p(rays), len(rays), _p(hits), mode,
                                threads or default_threads(), _p(nv), _p(nt))
        return (hits, nv, nt) if counts else hits

    def any(self, rays, brute=False, threads=None):
        """Occlusion (any triangle with t in [t_min, t_max)): the reference's
        traversal, or all triangles (brute)."""
        rays = np.ascontiguousarray(rays, np.float32)
        occ = np.zeros(len(rays), np.uint32)
        lib().orc_trace_any(self.h, _p(rays), len(rays), _p(occ), 1 if brute else 0, threads or default_threads())
        return occ

    def classify(self, rays, threads=None):
        rays = np.ascontiguousarray(rays, np.float32)
        q = np.zeros(len(rays), np.uint32)
        lib().orc_classify(self.h, _p(rays), len(rays), _p(q), threads or default_threads())
        return q

    def bvh4_closest(self, nodes, order, rays, threads=None):
        """Walk the core's BVH4 in the GPU kernel's order: hits and per-ray
        node / triangle visit counts and peak stack depth."""
        nodes = np.ascontiguousarray(nodes, np.uint32)
        order = np.ascontiguousarray(order, np.uint32)
        rays = np.ascontiguousarray(rays, np.float32)
        n = len(rays)
        hits = np.zeros((n, 4), np.uint32)
        nv, nt, ms = (np.zeros(n, np.uint32) for _ in range(3))
        lib().orc_bvh4_closest(self.h, _p(nodes), _p(order), _p(rays), n, _p(hits), _p(nv), _p(nt), _p(ms),
                               threads or default_threads())
        return hits, nv, nt, ms

    def bvh8q_closest(self, qnodes, order, rays, threads=None):
        """Closest hit over the core's 8-wide compressed nodes in the GPU
        kernel's order: hits (prim, t, u, v bits) and per-ray node / triangle
        visit counts and peak stack depth."""
        qnodes = np.ascontiguousarray(qnodes, np.uint32)
        order = np.ascontiguousarray(order, np.uint32)
        rays = np.ascontiguousarray(rays, np.float32)
        n = len(rays)
        hits = np.zeros((n, 4), np.uint32)
        nv, nt, ms = (np.zeros(n, np.uint32) for _ in range(3))
        lib().orc_bvh8q_closest(self.h, _p(qnodes), _p(order), _p(rays), n, _p(hits), _p(nv), _p(nt), _p(ms),
                                threads or default_threads())
        return hits, nv, nt, ms

    def bvh4q_any(self, qnodes, order, rays, threads=None):
        """Any hit over the core's quantized BVH4 in the GPU kernel's order:
        occlusion and per-ray node / triangle visit counts."""
        qnodes = np.ascontiguousarray(qnodes, np.uint32)
        order = np.ascontiguousarray(order, np.uint32)
        rays = np.ascontiguousarray(rays, np.float32)
        n = len(rays)
        occ, nv, nt = (np.zeros(n, np.uint32) for _ in range(3))
        lib().orc_bvh4q_any(self.h, _p(qnodes), _p(order), _p(rays), n, _p(occ), _p(nv), _p(nt),
                            threads or default_threads())
        return occ, nv, nt

    def bvh8q_any(self, qnodes, order, rays, threads=None):
        """Any hit over the core's 8-wide quantized BVH in the GPU kernel's
        order: occlusion, per-ray node / triangle visits and peak stack depth."""
        qnodes = np.ascontiguousarray(qnodes, np.uint32)
        order = np.ascontiguousarray(order, np.uint32)
        rays = np.ascontiguousarray(rays, np.float32)
        n = len(rays)
        occ, nv, nt, ms = (np.zeros(n, np.uint32) for _ in range(4))
        lib().orc_bvh8q_any(self.h, _p(qnodes), _p(order), _p(rays), n, _p(occ), _p(nv), _p(nt), _p(ms),
                            threads or default_threads())
        return occ, nv, nt, ms

    def alias_table(self):
        bins = np.zeros(self.info["lights"], BIN_DTYPE)
        lib().orc_scene_alias(self.h, _p(bins))
        return bins

    def bvh(self):
        nodes = np.zeros((self.info["nodes"], 10), np.float32)
        order = np.zeros(self.info["triangles"], np.uint32)
        lib().orc_scene_bvh(self.h, _p(nodes), _p(order))
        return nodes, order

    def render(self, settings, view, pixel_begin=0, pixel_end=None, threads=None, tprim=False):
        """HDR (accumulated radiance / spp) of pixels [pixel_begin, pixel_end).
        tprim: closest hits under the (t, prim id) order, the core's rule,
        instead of the reference's traversal order (they differ only where
        two triangles are hit at exactly the same t)."""
        w, h = settings["width"], settings["height"]
        pixel_end = w * h if pixel_end is None else pixel_end
        s = np.array([w, h, settings["samples_per_pixel"], settings["path_depth"], settings.get("use_nee", 1),
                      settings.get("rr_start_depth", 3), 1 if tprim else 0], np.uint32)
        cam = camera_basis(view, w / h)
        out = np.zeros((pixel_end - pixel_begin, 4), np.float32)
        cnt = np.zeros(2, np.uint64)
        lib().orc_render(self.h, _p(s), settings.get("seed", 1), _p(cam), pixel_begin, pixel_end,
                         threads or default_threads(), _p(out), _p(cnt))
        return out, int(cnt[0]), int(cnt[1])

    def trace_path(self, settings, view, sample, tprim=True):
        """Every ray of one sample's path: (records, radiance).  records is a
        structured array: kind (0 extension, 1 shadow), ray (pt_ray layout:
        origin, t_min, direction, t_max), result (extension: prim id bits, t,
        u, v; shadow: occluded, radiance before the clamp), depth."""
        w, h = settings["width"], settings["height"]
        s = np.array([w, h, settings["samples_per_pixel"], settings["path_depth"], settings.get("use_nee", 1),
                      settings.get("rr_start_depth", 3), 1 if tprim else 0], np.uint32)
        cam = camera_basis(view, w / h)
        rec = np.zeros((256, 16), np.float32)
        L = np.zeros(3, np.float32)
        n = lib().orc_trace_path(self.h, _p(s), settings.get("seed", 1), _p(cam), sample, _p(rec), 256, _p(L))
        rec = rec[:min(n, 256)]
        out = np.zeros(len(rec), [("kind", "<i4"), ("ray", "<f4", 8), ("result", "<f4", 4), ("depth", "<i4")])
        out["kind"] = rec[:, 0].astype(np.int32)
        out["ray"] = rec[:, 1:9]
        out["result"] = rec[:, 9:13]
        out["depth"] = rec[:, 13].astype(np.int32)
        return out, L


def camera_basis(view, aspect):
    pos = np.asarray(view.position, np.float32)
    q = np.asarray(view.rotation, np.float32)
    out = np.zeros(12, np.float32)
    lib().orc_camera_basis(C.c_float(view.fov_y), C.c_float(aspect), _p(pos), _p(q), _p(out))
    return out


def alias_build(weights):
    w = np.ascontiguousarray(weights, np.float32)
    bins = np.zeros(len(w), BIN_DTYPE)
    lib().orc_alias_build(_p(w), len(w), _p(bins))
    return bins


def alias_sample(bins, u):
    bins = np.ascontiguousarray(bins)
    u = np.ascontiguousarray(u, np.float32)
    idx = np.zeros(len(u), np.uint32)
    pmf = np.zeros(len(u), np.float32)
    lib().orc_alias_sample(_p(bins), len(bins), _p(u), len(u), _p(idx), _p(pmf))
    return idx, pmf


def pcg32(initstate, initseq, n, advance=0):
    out = np.zeros(n, np.uint32)
    lib().orc_pcg32(initstate, initseq, advance, n, _p(out))
    return out


def tonemap(hdr):
    hdr = np.ascontiguousarray(hdr, np.float32).reshape(-1, 4)
    out = np.zeros(len(hdr), np.uint32)
    lib().orc_tonemap(_p(hdr), len(hdr), _p(out))
    return out


REF_MIKK = os.path.join(HERE, "_ref", "libmikk_ref.so")
_mikk = None


def reference_tangents(positions, normals, uvs, indices):
    """The reference's own mikktspace (External/mikktspace/mikktspace.c, built
    by oracle/build.py into oracle/_ref) driven as the reference's import
    drives it (TangentSpace.cpp:14-89): (V, 4) float32 per-vertex tangents."""
    global _mikk
    if _mikk is None:
        if not os.path.exists(REF_MIKK):
            raise RuntimeError(f"{REF_MIKK} missing: run `python oracle/build.py` where the reference is present")
        L = C.CDLL(REF_MIKK)
        L.mikk_ref_tangents.restype = C.c_int
        L.mikk_ref_tangents.argtypes = [_P, _P, _P, _U32, _P, _U32, _P]
        _mikk = L
    P = np.ascontiguousarray(positions, np.float32).reshape(-1, 3)
    N = np.ascontiguousarray(normals, np.float32).reshape(-1, 3)
    U = np.ascontiguousarray(uvs, np.float32).reshape(-1, 2)
    idx = np.ascontiguousarray(indices, np.uint32).reshape(-1)
    out = np.zeros((len(P), 4), np.float32)
    _mikk.mikk_ref_tangents(_p(P), _p(N), _p(U), len(P), _p(idx), len(idx), _p(out))
    return out


# ---- shading functions of the oracle (the ones its render calls)

def _f32(a, cols=None):
    a = np.ascontiguousarray(a, np.float32)
    return a if cols is None else a.reshape(-1, cols)


def sincos(x):
    """(sin x, cos x) of sampling angles 0 <= x < 2 pi, as the render evaluates them."""
    x = _f32(x).ravel()
    s, c = np.zeros_like(x), np.zeros_like(x)
    lib().orc_sincos(_p(x), len(x), _p(s), _p(c))
    return s, c


def cosine_hemisphere(u2):
    u2 = _f32(u2, 2)
    w = np.zeros((len(u2), 3), np.float32)
    lib().orc_cosine_hemisphere(_p(u2), len(u2), _p(w))
    return w


def ggx_term(what, alpha, wm):
    """what: "D", "lambda" or "G1" of local directions wm (n, 3)."""
    wm = _f32(wm, 3)
    out = np.zeros(len(wm), np.float32)
    lib().orc_ggx_term({"D": 0, "lambda": 1, "G1": 2}[what], C.c_float(alpha), _p(wm), len(wm), _p(out))
    return out


def ggx_dvis(alpha, wo, wm):
    wo, wm = _f32(wo, 3), _f32(wm, 3)
    out = np.zeros(len(wm), np.float32)
    lib().orc_ggx_dvis(C.c_float(alpha), _p(wo), _p(wm), len(wm), _p(out))
    return out


def ggx_sample_wm(alpha, wo, u2):
    wo, u2 = _f32(wo, 3), _f32(u2, 2)
    wm = np.zeros((len(u2), 3), np.float32)
    lib().orc_ggx_sample_wm(C.c_float(alpha), _p(wo), _p(u2), len(u2), _p(wm))
    return wm


def _mat6(color, roughness, metallic, ior):
    return np.array(list(color) + [roughness, metallic, ior], np.float32)


def ggx_bsdf(material, wo, wl):
    """f(wo, wl) of the NEE path and whether the sample is accepted; material =
    (color, roughness, metallic, ior)."""
    m = _mat6(*material)
    wo, wl = _f32(wo, 3), _f32(wl, 3)
    f = np.zeros((len(wl), 3), np.float32)
    ok = np.zeros(len(wl), np.uint32)
    lib().orc_ggx_bsdf(_p(m), _p(wo), _p(wl), len(wl), _p(f), _p(ok))
    return f, ok.astype(bool)


def ggx_pdf(material, wo, wl):
    m = _mat6(*material)
    wo, wl = _f32(wo, 3), _f32(wl, 3)
    pdf = np.zeros(len(wl), np.float32)
    lib().orc_ggx_pdf(_p(m), _p(wo), _p(wl), len(wl), _p(pdf))
    return pdf


def ggx_sample(material, wo, u3):
    """The indirect GGX sample: (wl, f, weight, pdf, accepted)."""
    m = _mat6(*material)
    wo, u3 = _f32(wo, 3), _f32(u3, 3)
    n = len(u3)
    wl, f, wt = (np.zeros((n, 3), np.float32) for _ in range(3))
    pdf = np.zeros(n, np.float32)
    ok = np.zeros(n, np.uint32)
    lib().orc_ggx_sample(_p(m), _p(wo), _p(u3), n, _p(wl), _p(f), _p(wt), _p(pdf), _p(ok))
    return wl, f, wt, pdf, ok.astype(bool)
