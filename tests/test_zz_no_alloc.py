"""The allocation-free closest-hit trace against the allocating one, through the C ABI.

Device buffers come from the HIP runtime the library itself loaded
(libamdhip64, through ctypes)."""
import ctypes as C

import numpy as np
import pytest

import ptcore
import scenes

pytestmark = pytest.mark.gpu

SENTINEL = 0x7F7F7F7F
H2D, D2H = 1, 2


def _hip():
    hip = C.CDLL("libamdhip64.so")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipFree.argtypes = [C.c_void_p]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipDeviceSynchronize.argtypes = []
    hip.hipMemGetInfo.argtypes = [C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]
    return hip


class _DeviceArray:
    def __init__(self, hip, host):
        self.hip, self.bytes = hip, host.nbytes
        self.ptr = C.c_void_p()
        assert hip.hipMalloc(C.byref(self.ptr), self.bytes) == 0
        assert hip.hipMemcpy(self.ptr, host.ctypes.data, self.bytes, H2D) == 0

    def read(self, like):
        out = np.empty_like(like)
        assert self.hip.hipMemcpy(out.ctypes.data, self.ptr, self.bytes, D2H) == 0
        return out

    def free(self):
        self.hip.hipFree(self.ptr)


def _trace(hip, r, fn, d_rays, n):
    init = np.full((n + 256, 4), SENTINEL, np.uint32)
    out = _DeviceArray(hip, init)
    try:
        st = getattr(r.lib, fn)(r.h, d_rays.ptr, n, out.ptr, None)
        assert hip.hipDeviceSynchronize() == 0
        assert st == ptcore.PT_OK, (fn, st, r.lib.pt_get_error_message(r.h))
        return out.read(init)
    finally:
        out.free()


def test_no_alloc_trace_equals_the_device_trace(oracle_lib):
    """Production scene (983,042 triangles, a BVH whose stack spills past LDS):
    two calls in a row, 2^20 rays then 2^16, each equal bit for bit to
    pt_trace_closest_device on the same rays, nothing written past the batch,
    and a subset equal to the oracle's (t, prim id) minimum."""
    s = scenes.production_scene()
    o = oracle_lib.OracleScene(s)
    with ptcore.Renderer() as r:
        r.InitializeRenderingBuffers(64, 64)
        r.InitializeSceneBuffers(s)
        info = r.bvh_info()
        assert info["stack_total"] > info["stack_lds"]
        hip = _hip()
        free, total = C.c_size_t(), C.c_size_t()
        hip.hipMemGetInfo(C.byref(free), C.byref(total))
        print(f"device memory free {free.value / 2**30:.1f} of {total.value / 2**30:.1f} GiB")
        for k, n in enumerate((1 << 20, 1 << 16)):
            rays = scenes.random_rays(n, (-10, -10, -10), (110, 110, 110), 400 + k)
            d_rays = _DeviceArray(hip, rays)
            try:
                want = _trace(hip, r, "pt_trace_closest_device", d_rays, n)
                got = _trace(hip, r, "pt_trace_closest_device_noalloc", d_rays, n)
            finally:
                d_rays.free()
            assert np.array_equal(got[:n], want[:n])
            assert (got[n:] == SENTINEL).all() and (want[n:] == SENTINEL).all()
            sub = np.arange(0, n, 97)
            assert np.array_equal(got[sub], o.closest(rays[sub], tprim=True))
            assert (got[:n, 0] != ptcore.INVALID).mean() > 0.01
