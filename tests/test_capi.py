"""CPU checks of the C ABI: the library loads, exports every symbol that
include/pt_core.h declares, and the Python binding covers them.  No compute."""
import ctypes
import os
import re

import pytest

import ptcore

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "pt_core.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pt_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    for must in ("pt_initialize_scene_buffers", "pt_initialize_rendering_buffers", "pt_start_simulation",
                 "pt_render_frame", "pt_trace_closest", "pt_trace_any", "pt_alias_sample", "pt_classify_hits"):
        assert must in syms


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(ptcore.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_binding_covers_every_declared_symbol():
    assert sorted(ptcore.SIGNATURES) == declared_symbols()
    ptcore.load_library()


def test_version_string():
    lib = ptcore.load_library()
    assert b"gfx950" in lib.pt_version()


def test_null_handles_are_rejected():
    lib = ptcore.load_library()
    assert lib.pt_initialize_rendering_buffers(None, 4, 4) == 1
    assert lib.pt_is_rendering(None) == 0
    assert lib.pt_renderer_create(None) == 1


def test_create_without_gpu_reports_an_error_status():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    lib = ptcore.load_library()
    h = ctypes.c_void_p()
    st = lib.pt_renderer_create(ctypes.byref(h))
    assert st == 4   # PT_ERROR_DEVICE: no HIP device is visible
    assert not h.value


def test_cpp_facade_compiles_and_links(tmp_path):
    """The header-only C++ facade (reference method names) builds against the C ABI."""
    import subprocess
    src = tmp_path / "use_facade.cpp"
    src.write_text(r'''
#include <pt/Renderer.hpp>
#include <cstdio>
int main() {
    pt::Renderer r;
    if (!r.Created()) { std::printf("no device: %s\n", r.Created().message.c_str()); return 0; }
    auto ok = r.InitializeRenderingBuffers(16, 16);
    std::printf("%d %u\n", (int)ok.status, r.GetPixelGridWidth());
    return 0;
}
''')
    inc = os.path.join(ROOT, "gpu-path-tracer_amd", "include")
    libdir = os.path.dirname(ptcore.LIB_PATH)
    exe = tmp_path / "use_facade"
    r = subprocess.run(["g++", "-std=c++17", "-I", inc, str(src), "-o", str(exe), f"-L{libdir}", "-lptcore",
                        f"-Wl,-rpath,{libdir}"], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stderr
