"""OBJ/MTL import (pt_import_obj, host code of libptcore.so): geometry,
the reference's material conversion (sRGB u8 quantisation of Kd and Ks, the
`surface` key, default RMA and normal-map texels, fp32 emission), polygon
triangulation, relative indices, required normals and uvs, default material,
tangents equal to the reference's mikktspace; and on the GPU a render of an
imported room against the oracle."""
import math
import os
import sys

import numpy as np
import pytest

import ptcore
import scenes

SURFACES = {"diffuse": 1, "lambert": 1, "unlit": 1, "mirror": 2, "phong": 3, "ggx": 4, "microfacet": 4,
            "emissive": 5, "normal": 0, "nonsense": 1}


def _srgb_roundtrip(x):
    """Manager.cpp:64-79 encode to u8 (lround), then the sRGB decode."""
    x = np.clip(np.asarray(x, np.float64), 0.0, 1.0)
    s = np.where(x <= 0.0031308, 12.92 * x, 1.055 * np.power(x, 1.0 / 2.4) - 0.055)
    u = np.floor(s * 255.0 + 0.5)
    c = u / 255.0
    return np.where(c <= 0.04045, c / 12.92, np.power((c + 0.055) / 1.055, 2.4))


def _write_scene_obj(scene, path, names):
    """The scene as OBJ + MTL: every triangle with its own corners, the flat
    normals of the scene and uvs (0,0), (1,0), (0,1)."""
    lines = ["mtllib room.mtl"]
    P, N = scene.positions, scene.normals
    for t in range(len(P)):
        for k in range(3):
            lines.append("v %.9g %.9g %.9g" % tuple(P[t, k]))
            lines.append("vn %.9g %.9g %.9g" % tuple(N[t, k]))
    lines += ["vt 0 0", "vt 1 0", "vt 0 1"]
    cur = None
    for t in range(len(P)):
        m = int(scene.material_ids[t])
        if m != cur:
            lines.append(f"usemtl {names[m]}")
            cur = m
        a = 3 * t + 1
        lines.append(f"f {a}/1/{a} {a + 1}/2/{a + 1} {a + 2}/3/{a + 2}")
    (path / "room.obj").write_text("\n".join(lines) + "\n")
    mtl = []
    for m, name in enumerate(names):
        rec = scene.materials[m]
        surf = {1: "diffuse", 2: "mirror", 4: "ggx", 5: "emissive"}[int(rec["model"])]
        mtl += [f"newmtl {name}", "Kd %.9g %.9g %.9g" % tuple(rec["color"]),
                "Ke %.9g %.9g %.9g" % tuple(rec["emission"]), f"surface {surf}", ""]
    (path / "room.mtl").write_text("\n".join(mtl))
    return path / "room.obj"


def test_import_cornell_room(tmp_path):
    s = scenes.cornell_box()
    names = ["white", "red", "green", "light"]
    got = ptcore.import_obj(_write_scene_obj(s, tmp_path, names))
    assert got.triangle_count == 36
    # triangles regrouped by material (ascending), face order kept within a group
    order = np.argsort(s.material_ids, kind="stable")
    assert np.array_equal(got.positions, s.positions[order])
    assert np.allclose(got.normals, s.normals[order], atol=1e-6)
    assert np.array_equal(got.material_ids, s.material_ids[order])
    m = got.materials
    assert np.allclose(m["color"], _srgb_roundtrip(s.materials["color"]), rtol=1e-6, atol=1e-7)
    # the known value of the reference's quantisation: 0.73 -> u8 222 -> 0.7304
    assert abs(m["color"][0, 0] - 0.7304) < 1e-4
    assert np.array_equal(m["emission"], s.materials["emission"])
    assert np.array_equal(m["model"], s.materials["model"])
    assert (m["roughness"] == 1.0).all() and (m["metallic"] == 0.0).all()       # default RMA texel
    assert np.allclose(m["normal_ts"], [1.0 / 255.0, 1.0 / 255.0, 1.0], atol=1e-7)


def test_surface_keys_polygons_and_relative_indices(tmp_path):
    mtl = []
    for k in SURFACES:
        mtl += [f"newmtl m_{k}", "Kd 0.5 0.25 1", "Ks 0.2 0.2 0.2", "Ni 1.45", f"surface {k}", ""]
    mtl += ["newmtl m_none", "Kd 0.3 0.3 0.3", ""]
    (tmp_path / "lib.mtl").write_text("\n".join(mtl))
    obj = ["mtllib lib.mtl", "v 0 0 0", "v 1 0 0", "v 1 1 0", "v 0 1 0", "v 0.5 1.5 0",
           "vt 0 0", "vt 1 0", "vt 1 1", "vt 0 1", "vt 0.5 1.5", "vn 0 0 2"]
    # a pentagon before any usemtl (default material), then a quad per material with relative indices
    obj.append("f 1/1/1 2/2/1 3/3/1 5/5/1 4/4/1")
    for k in list(SURFACES) + ["none"]:
        obj += [f"usemtl m_{k}", "f -5/-5/-1 -4/-4/-1 -3/-3/-1 -2/-2/-1"]
    (tmp_path / "a.obj").write_text("\n".join(obj) + "\n")
    s = ptcore.import_obj(tmp_path / "a.obj")
    nm = len(SURFACES) + 1
    assert len(s.materials) == nm + 1                       # the library, then the default material
    assert s.triangle_count == 3 + 2 * nm                   # fan: pentagon 3, quads 2
    default = s.materials[nm]
    assert default["model"] == 1 and np.allclose(default["color"], 1.0)
    for i, k in enumerate(SURFACES):
        assert s.materials[i]["model"] == SURFACES[k], k
        assert abs(s.materials[i]["ior"] - 1.45) < 1e-7
    assert s.materials[nm - 1]["model"] == 1                # no surface key: unlit, shaded as diffuse
    assert np.allclose(s.materials[0]["color"], _srgb_roundtrip([0.5, 0.25, 1.0]), atol=1e-7)
    assert np.allclose(s.materials[0]["specular"], _srgb_roundtrip([0.2, 0.2, 0.2]), atol=1e-7)
    # the default-material faces come first (none sorts before the library ids)
    assert (s.material_ids[:3] == nm).all()
    assert np.array_equal(s.positions[0], [[0, 0, 0], [1, 0, 0], [1, 1, 0]])
    # normals are normalised; uv axes along x and y give tangent +x, handedness +1
    assert np.allclose(s.normals, [0, 0, 1])
    assert np.allclose(s.tangents[..., :3], [1, 0, 0], atol=1e-6) and (s.tangents[..., 3] == 1).all()


@pytest.mark.parametrize("face, message", [("f 1//1 2//1 3//1", "uv data is required"),
                                           ("f 1/1 2/2 3/3", "normal data is required"),
                                           ("f 1/1/1 2/2/1 9/3/1", "vertex index 9 out of range"),
                                           ("f 1/1/1 2/2/1", "three corners")])
def test_import_errors(tmp_path, face, message):
    (tmp_path / "b.obj").write_text("v 0 0 0\nv 1 0 0\nv 0 1 0\nvt 0 0\nvt 1 0\nvt 0 1\nvn 0 0 1\n" + face + "\n")
    with pytest.raises(ptcore.PtError) as e:
        ptcore.import_obj(tmp_path / "b.obj")
    assert e.value.status == 7 and message in str(e.value)           # PT_ERROR_IO


def test_import_missing_file(tmp_path):
    with pytest.raises(ptcore.PtError) as e:
        ptcore.import_obj(tmp_path / "missing.obj")
    assert "cannot open" in str(e.value)


def test_imported_tangents_equal_reference_mikktspace(tmp_path):
    """A bumpy grid with mirrored uvs written as OBJ (one index per corner for
    v, vt and vn): the imported per-corner tangents equal the reference's
    mikktspace run over the same vertices, normals as imported."""
    sys.path.insert(0, os.path.join(os.path.dirname(__file__), "..", "oracle"))
    import build as oracle_build
    import oracle
    if not os.path.exists(oracle.REF_MIKK) and oracle_build.build_ref() is None:
        pytest.skip("the reference's mikktspace sources are not present and oracle/_ref is not built")
    n = 12
    x, y = np.meshgrid(np.linspace(0, 1, n + 1), np.linspace(0, 1, n + 1))
    z = 0.2 * np.sin(5 * x) * np.cos(4 * y)
    P = np.stack([x, y, z], -1).reshape(-1, 3).astype(np.float32)
    N = np.stack([-np.cos(5 * x) * np.cos(4 * y), 0.8 * np.sin(5 * x) * np.sin(4 * y), np.ones_like(x)], -1)
    N = N.reshape(-1, 3).astype(np.float32)                         # not unit: the import normalises
    U = np.stack([np.abs(x - 0.5), y], -1).reshape(-1, 2).astype(np.float32)
    i = np.arange(n)[:, None] * (n + 1) + np.arange(n)[None, :]
    q = np.stack([i, i + 1, i + n + 2, i + n + 1], -1).reshape(-1, 4)
    lines = ["v %.9g %.9g %.9g" % tuple(p) for p in P] + ["vt %.9g %.9g" % tuple(u) for u in U] + \
            ["vn %.9g %.9g %.9g" % tuple(v) for v in N]
    lines += ["f " + " ".join(f"{k + 1}/{k + 1}/{k + 1}" for k in quad) for quad in q]   # quads, fan-split
    (tmp_path / "g.obj").write_text("\n".join(lines) + "\n")
    s = ptcore.import_obj(tmp_path / "g.obj")
    idx = np.concatenate([q[:, [0, 1, 2]], q[:, [0, 2, 3]]], 1).reshape(-1)
    assert s.triangle_count == 2 * n * n
    Nv = np.zeros_like(P)
    Nv[idx] = s.normals.reshape(-1, 3)
    assert np.allclose(np.linalg.norm(Nv, axis=1), 1.0, atol=1e-6)
    want = oracle.reference_tangents(P, Nv, U, idx)[idx]
    assert np.array_equal(s.tangents.reshape(-1, 4), want)
    assert (want[:, 3] == 1).any() and (want[:, 3] == -1).any()    # both sides of the mirror


@pytest.mark.gpu
def test_imported_room_renders_like_the_oracle(tmp_path, gpu_renderer, oracle_lib):
    s = scenes.cornell_box()
    room = ptcore.import_obj(_write_scene_obj(s, tmp_path, ["white", "red", "green", "light"]))
    room.view = s.view
    gpu_renderer.InitializeRenderingBuffers(48, 48)
    gpu_renderer.InitializeSceneBuffers(room)
    st = dict(width=48, height=48, samples_per_pixel=64, path_depth=8, use_nee=1, rr_start_depth=3, seed=9)
    gpu_renderer.RenderFrame(st, room.view)
    hdr = gpu_renderer.GetFrameHdr(48, 48).reshape(-1, 4)
    ref, _, _ = oracle_lib.OracleScene(room).render(st, room.view)
    rm = float(np.mean((hdr[:, :3] - ref[:, :3]) ** 2 / (ref[:, :3] ** 2 + 1e-2)))
    assert rm < 1e-3, rm
    assert math.isfinite(float(hdr.sum()))
