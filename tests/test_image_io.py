"""Image files (pt_write_image, pt_save_frame): PNG read back by Pillow's
decoder, OpenEXR read back by a reader written from the published file layout
(magic and version, attribute header, scanline offset table, per-line chunks
with channels in name order), row order, and SaveImage's argument errors
(ImageIO.cpp:119-160)."""
import struct

import numpy as np
import pytest
from PIL import Image

import ptcore
import scenes


def _read_exr(path):
    """Single-part scanline OpenEXR without compression -> header dict and
    {channel: (H, W) array}."""
    b = open(path, "rb").read()
    assert b[:4] == bytes([0x76, 0x2F, 0x31, 0x01])
    assert struct.unpack_from("<I", b, 4)[0] == 2              # version 2, no flags: single-part scanline
    pos, header = 8, {}
    while b[pos] != 0:
        name_end = b.index(b"\0", pos)
        name = b[pos:name_end].decode()
        type_end = b.index(b"\0", name_end + 1)
        typ = b[name_end + 1:type_end].decode()
        size = struct.unpack_from("<i", b, type_end + 1)[0]
        header[name] = (typ, b[type_end + 5:type_end + 5 + size])
        pos = type_end + 5 + size
    pos += 1
    typ, raw = header["channels"]
    assert typ == "chlist"
    chans, q = [], 0
    while raw[q] != 0:
        e = raw.index(b"\0", q)
        ptype, _lin, xs, ys = struct.unpack_from("<iB3xii", raw, e + 1)
        chans.append((raw[q:e].decode(), ptype, xs, ys))
        q = e + 17
    assert header["compression"] == ("compression", b"\0")
    assert header["lineOrder"] == ("lineOrder", b"\0")
    x0, y0, x1, y1 = struct.unpack("<4i", header["dataWindow"][1])
    assert header["displayWindow"] == header["dataWindow"]
    assert struct.unpack("<f", header["pixelAspectRatio"][1]) == (1.0,)
    assert header["screenWindowCenter"][0] == "v2f" and header["screenWindowWidth"][0] == "float"
    w, h = x1 - x0 + 1, y1 - y0 + 1
    offsets = struct.unpack_from(f"<{h}Q", b, pos)
    out = {c[0]: np.zeros((h, w), np.float32) for c in chans}
    for k, off in enumerate(offsets):
        y, n = struct.unpack_from("<ii", b, off)
        assert y == y0 + k and n == 4 * w * len(chans)
        q = off + 8
        for name, ptype, xs, ys in chans:
            assert ptype == 2 and xs == 1 and ys == 1               # FLOAT, full resolution
            out[name][k] = np.frombuffer(b, "<f4", w, q)
            q += 4 * w
    return [c[0] for c in chans], out


def test_png_roundtrip_through_pillow(tmp_path):
    rng = np.random.default_rng(1)
    px = rng.integers(0, 2**32, (37, 53), dtype=np.uint64).astype(np.uint32)
    rgba = px.view(np.uint8).reshape(37, 53, 4)
    ptcore.write_image(tmp_path / "a.png", px)
    with Image.open(tmp_path / "a.png") as im:
        assert im.mode == "RGBA" and im.size == (53, 37)
        assert np.array_equal(np.asarray(im), rgba)
    ptcore.write_image(tmp_path / "b.PNG", px, bottom_up=True)        # frame order: row 0 is the bottom
    with Image.open(tmp_path / "b.PNG") as im:
        assert np.array_equal(np.asarray(im), rgba[::-1])


def test_exr_roundtrip(tmp_path):
    rng = np.random.default_rng(2)
    hdr = rng.normal(0, 10, (19, 31, 4)).astype(np.float32)
    hdr[0, 0] = (0.0, -0.0, 1e30, 1e-40)                             # zero signs, huge, subnormal
    ptcore.write_image(tmp_path / "a.exr", hdr)
    names, ch = _read_exr(tmp_path / "a.exr")
    assert names == ["A", "B", "G", "R"]
    got = np.stack([ch["R"], ch["G"], ch["B"], ch["A"]], -1)
    assert got.view(np.uint32).tobytes() == hdr.view(np.uint32).tobytes()
    ptcore.write_image(tmp_path / "b.exr", hdr, bottom_up=True)
    _, ch = _read_exr(tmp_path / "b.exr")
    assert np.array_equal(np.stack([ch["R"], ch["G"], ch["B"], ch["A"]], -1), hdr[::-1])


def test_exr_header_bytes(tmp_path):
    """The channel list entry layout: name, pixel type 2 (FLOAT), pLinear and
    three reserved bytes, x and y sampling; 4 entries and a terminator."""
    ptcore.write_image(tmp_path / "c.exr", np.ones((2, 3, 4), np.float32))
    b = open(tmp_path / "c.exr", "rb").read()
    entry = lambda c: c + b"\0" + struct.pack("<i", 2) + b"\0\0\0\0" + struct.pack("<ii", 1, 1)
    chl = b"".join(entry(c) for c in (b"A", b"B", b"G", b"R")) + b"\0"
    assert b[8:8 + 9 + 7 + 4 + len(chl)] == b"channels\0chlist\0" + struct.pack("<i", len(chl)) + chl
    assert b"dataWindow\0box2i\0" + struct.pack("<5i", 16, 0, 0, 2, 1) in b


@pytest.mark.parametrize("name, pixels, status", [
    ("", np.zeros((2, 2), np.uint32), 1),                 # empty path
    ("noext", np.zeros((2, 2), np.uint32), 1),            # missing extension
    ("x.png", np.zeros((0, 2), np.uint32), 1),            # zero size
    ("x.jpg", np.zeros((2, 2), np.uint32), 6),            # no writer
    ("x.png", np.zeros((2, 2, 4), np.float32), 6),        # PNG takes RGBA8
    ("x.exr", np.zeros((2, 2), np.uint32), 6),            # EXR takes float RGBA
    ("missing_dir_is_a_file/x.png", np.zeros((2, 2), np.uint32), 7),
])
def test_write_image_errors(tmp_path, name, pixels, status):
    (tmp_path / "missing_dir_is_a_file").write_text("")
    path = (tmp_path / name) if name else ""
    with pytest.raises(ptcore.PtError) as e:
        ptcore.write_image(path, pixels)
    assert e.value.status == status


@pytest.mark.gpu
def test_save_frame_png_and_exr(tmp_path, gpu_renderer):
    s = scenes.cornell_box()
    r = gpu_renderer
    r.InitializeRenderingBuffers(40, 24)
    r.InitializeSceneBuffers(s)
    r.RenderFrame(dict(width=40, height=24, samples_per_pixel=8, path_depth=4, use_nee=1, rr_start_depth=3,
                       seed=5), s.view)
    frame = r.GetFrame(40, 24)
    hdr = r.GetFrameHdr(40, 24)
    r.SaveFrame(tmp_path / "f.png")
    r.SaveFrame(tmp_path / "f.exr")
    with Image.open(tmp_path / "f.png") as im:
        assert np.array_equal(np.asarray(im), frame.view(np.uint8).reshape(24, 40, 4)[::-1])
    _, ch = _read_exr(tmp_path / "f.exr")
    assert np.array_equal(np.stack([ch["R"], ch["G"], ch["B"], ch["A"]], -1), hdr[::-1])
    with pytest.raises(ptcore.PtError) as e:
        r.SaveFrame(tmp_path / "f.tiff")
    assert e.value.status == 6
