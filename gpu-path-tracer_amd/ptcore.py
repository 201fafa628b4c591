"""ctypes binding of libptcore.so (include/pt_core.h) and a Python mirror of
the reference Renderer's interface
(Core/Include/Core/Graphics/Cuda/PathTracing/Renderer.hpp:42-189).

The library is the product: there is no fallback.  Loading fails loudly if
libptcore.so has not been built.
"""
import contextlib
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# PTCORE_LIB selects another build of the same library (variant timing)
LIB_PATH = os.environ.get("PTCORE_LIB", os.path.join(HERE, "libptcore.so"))

PT_OK = 0
STATUS_NAMES = {0: "PT_OK", 1: "PT_ERROR_INVALID_ARGUMENT", 2: "PT_ERROR_INVALID_STATE", 3: "PT_ERROR_OUT_OF_RANGE",
                4: "PT_ERROR_DEVICE", 5: "PT_ERROR_OUT_OF_MEMORY", 6: "PT_ERROR_UNSUPPORTED", 7: "PT_ERROR_IO"}
QUEUE_DIFFUSE, QUEUE_GGX, QUEUE_MIRROR, QUEUE_EMISSIVE, QUEUE_REGEN = range(5)
INVALID = 0xFFFFFFFF
REQUEST_QUEUED, REQUEST_RUNNING, REQUEST_DONE, REQUEST_FAILED = 0, 1, 2, 3
BVH_BUILDER_HOST, BVH_BUILDER_GPU = 0, 1
FRAME_DEFAULT, FRAME_PROFILE, FRAME_INSTRUMENT, FRAME_NO_GRAPH = 0, 1, 2, 4


class PtError(RuntimeError):
    def __init__(self, status, message):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {message}")
        self.status = status


class SceneDesc(C.Structure):
    _fields_ = [("triangle_count", C.c_uint32), ("positions", C.c_void_p), ("normals", C.c_void_p),
                ("tangents", C.c_void_p), ("material_ids", C.c_void_p), ("material_count", C.c_uint32),
                ("materials", C.c_void_p), ("light_weights", C.c_void_p)]


class RenderSettings(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("samples_per_pixel", C.c_uint32),
                ("path_depth", C.c_uint32), ("use_nee", C.c_uint32), ("rr_start_depth", C.c_uint32),
                ("seed", C.c_uint64)]


class Camera(C.Structure):
    _fields_ = [("fov_y", C.c_float), ("near_plane", C.c_float), ("far_plane", C.c_float)]


class MotionState(C.Structure):
    _fields_ = [("position", C.c_float * 3), ("rotation", C.c_float * 4)]


class View(C.Structure):
    _fields_ = [("scene", C.c_uint32), ("reserved", C.c_uint32), ("settings", RenderSettings), ("camera", Camera),
                ("state", MotionState)]


class ServerConfig(C.Structure):
    _fields_ = [("accum_budget_bytes", C.c_uint64), ("window_samples", C.c_uint64), ("chunk_samples", C.c_uint64),
                ("max_requests", C.c_uint32), ("reserved", C.c_uint32)]


class RequestInfo(C.Structure):
    _fields_ = [("id", C.c_uint64), ("state", C.c_uint32), ("scene", C.c_uint32), ("samples", C.c_uint64),
                ("submit_ms", C.c_double), ("admit_ms", C.c_double), ("done_ms", C.c_double)]


class ServerStats(C.Structure):
    _fields_ = [("requests_submitted", C.c_uint64), ("requests_done", C.c_uint64), ("requests_failed", C.c_uint64),
                ("samples", C.c_uint64), ("extension_rays", C.c_uint64), ("shadow_rays", C.c_uint64),
                ("iterations", C.c_uint64), ("wall_ms", C.c_double), ("busy_ms", C.c_double)]


class GuideSettings(C.Structure):
    _fields_ = [("passes", C.c_uint32), ("train_steps", C.c_uint32), ("batch", C.c_uint32),
                ("learning_rate", C.c_float), ("guide_fraction", C.c_float), ("weight_max", C.c_float),
                ("record_capacity", C.c_uint32), ("reset", C.c_uint32), ("seed", C.c_uint64)]


class GuideStats(C.Structure):
    _fields_ = [("render_ms", C.c_double), ("train_ms", C.c_double), ("records", C.c_uint64),
                ("loss_first", C.c_float), ("loss_last", C.c_float), ("steps", C.c_uint32), ("passes", C.c_uint32),
                ("train_kernel_us", C.c_double), ("shaded_entries", C.c_uint64)]


GUIDE_PARAMS = 16640
GUIDE_DIM = 64
GUIDE_RECORD_DTYPE = np.dtype([("position", "<f4", 3), ("prev", "<u4"), ("normal", "<f4", 3), ("pdf", "<f4"),
                               ("wo", "<f4", 3), ("cos_i", "<f4"), ("wi", "<f4", 3), ("pad0", "<f4"),
                               ("throughput", "<f4", 3), ("pad1", "<f4"), ("radiance", "<f4", 3), ("pad2", "<f4")])
assert GUIDE_RECORD_DTYPE.itemsize == 96


class Progress(C.Structure):
    _fields_ = [("is_rendering", C.c_uint32), ("done_frames", C.c_uint32), ("total_frames", C.c_uint32),
                ("width", C.c_uint32), ("height", C.c_uint32), ("samples_per_pixel", C.c_uint32),
                ("samples_per_pixel_axis", C.c_uint32), ("path_depth", C.c_uint32),
                ("done_samples", C.c_uint64), ("total_samples", C.c_uint64)]


class FrameStats(C.Structure):
    _fields_ = [("render_ms", C.c_double), ("extension_rays", C.c_uint64), ("shadow_rays", C.c_uint64),
                ("samples", C.c_uint64), ("iterations", C.c_uint32), ("closest_launches", C.c_uint32),
                ("closest_ms", C.c_double), ("shadow_ms", C.c_double), ("shade_ms", C.c_double),
                ("regen_ms", C.c_double), ("closest_nodes", C.c_uint64), ("closest_tris", C.c_uint64),
                ("shadow_nodes", C.c_uint64), ("shadow_tris", C.c_uint64), ("bin_ms", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class AliasBin(C.Structure):
    _fields_ = [("q", C.c_float), ("p", C.c_float), ("alias", C.c_uint32), ("value", C.c_uint32)]


ALIAS_DTYPE = np.dtype([("q", "<f4"), ("p", "<f4"), ("alias", "<u4"), ("value", "<u4")])


class BvhInfo(C.Structure):
    _fields_ = [("node_count", C.c_uint32), ("leaf_count", C.c_uint32), ("max_depth", C.c_uint32),
                ("max_leaf_size", C.c_uint32), ("triangle_count", C.c_uint32), ("light_count", C.c_uint32),
                ("stack_lds", C.c_uint32), ("stack_total", C.c_uint32), ("node_width", C.c_uint32),
                ("bvh8_node_count", C.c_uint32), ("bvh8_depth", C.c_uint32)]


class QueueOccupancy(C.Structure):
    _fields_ = [("extension", C.c_uint32), ("material", C.c_uint32 * 4), ("missed", C.c_uint32),
                ("shadow", C.c_uint32), ("next", C.c_uint32), ("capacity", C.c_uint32)]


QUEUE_OCCUPANCY_DTYPE = np.dtype([("extension", "<u4"), ("material", "<u4", 4), ("missed", "<u4"),
                                  ("shadow", "<u4"), ("next", "<u4"), ("capacity", "<u4")])


class Tuning(C.Structure):
    _fields_ = [("trace_block", C.c_uint32), ("trace_blocks_per_cu", C.c_uint32), ("shade_block", C.c_uint32),
                ("pool_size", C.c_uint32), ("use_graph", C.c_uint32), ("bvh_max_leaf", C.c_uint32),
                ("bvh_traversal_cost", C.c_float), ("trace_refill", C.c_uint32), ("bvh_width", C.c_uint32),
                ("bvh_builder", C.c_uint32), ("bvh_quantized", C.c_uint32), ("shadow_blocks_per_cu", C.c_uint32),
                ("stack_limit", C.c_uint32)]


_P = C.c_void_p
_U32 = C.c_uint32
# name: (restype, argtypes); every symbol of include/pt_core.h
SIGNATURES = {
    "pt_version": (C.c_char_p, []),
    "pt_renderer_create": (C.c_int, [C.POINTER(_P)]),
    "pt_renderer_destroy": (None, [_P]),
    "pt_get_error_message": (C.c_char_p, [_P]),
    "pt_initialize_rendering_buffers": (C.c_int, [_P, _U32, _U32]),
    "pt_initialize_scene_buffers": (C.c_int, [_P, C.POINTER(SceneDesc)]),
    "pt_clear": (C.c_int, [_P, C.c_uint8, C.c_uint8, C.c_uint8, C.c_uint8]),
    "pt_start_simulation": (C.c_int, [_P, C.POINTER(RenderSettings), C.POINTER(Camera), C.POINTER(MotionState), _U32,
                                      C.c_char_p]),
    "pt_resume_simulation": (C.c_int, [_P, _U32]),
    "pt_stop_rendering": (C.c_int, [_P]),
    "pt_is_rendering": (C.c_int, [_P]),
    "pt_get_progress": (C.c_int, [_P, C.POINTER(Progress)]),
    "pt_frame_copy": (C.c_int, [_P, _P, C.c_size_t]),
    "pt_frame_copy_hdr": (C.c_int, [_P, _P, C.c_size_t]),
    "pt_frame_lock": (C.c_int, [_P, C.POINTER(C.POINTER(C.c_uint32)), C.POINTER(C.c_size_t)]),
    "pt_frame_unlock": (C.c_int, [_P]),
    "pt_peek_last_error": (C.c_size_t, [_P, C.c_char_p, C.c_size_t]),
    "pt_consume_last_error": (C.c_size_t, [_P, C.c_char_p, C.c_size_t]),
    "pt_render_frame": (C.c_int, [_P, C.POINTER(RenderSettings), C.POINTER(Camera), C.POINTER(MotionState), _U32,
                                  C.POINTER(FrameStats)]),
    "pt_add_scene": (C.c_int, [_P, C.POINTER(SceneDesc), C.POINTER(_U32)]),
    "pt_remove_scene": (C.c_int, [_P, _U32]),
    "pt_render_batch": (C.c_int, [_P, C.POINTER(View), _U32, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p), _U32,
                                  C.POINTER(FrameStats)]),
    "pt_server_start": (C.c_int, [_P, C.POINTER(ServerConfig)]),
    "pt_server_submit": (C.c_int, [_P, C.POINTER(View), _P, _P, C.POINTER(C.c_uint64)]),
    "pt_server_wait": (C.c_int, [_P, C.c_uint64, _U32, C.POINTER(RequestInfo)]),
    "pt_server_get_stats": (C.c_int, [_P, C.POINTER(ServerStats)]),
    "pt_server_stop": (C.c_int, [_P, C.POINTER(ServerStats)]),
    "pt_render_guided": (C.c_int, [_P, C.POINTER(RenderSettings), C.POINTER(Camera), C.POINTER(MotionState),
                                   C.POINTER(GuideSettings), C.POINTER(GuideStats), C.POINTER(FrameStats)]),
    "pt_guide_reset": (C.c_int, [_P, C.c_uint64]),
    "pt_guide_get_weights": (C.c_int, [_P, _P, _U32]),
    "pt_guide_set_weights": (C.c_int, [_P, _P, _U32]),
    "pt_guide_set_box": (C.c_int, [_P, _P, _P]),
    "pt_guide_forward": (C.c_int, [_P, _P, _U32, _P]),
    "pt_guide_mixture": (C.c_int, [_P, _P, _U32, _P]),
    "pt_guide_gradient": (C.c_int, [_P, _P, _U32, C.c_float, _P, _P]),
    "pt_guide_train": (C.c_int, [_P, _P, _U32, _U32, _U32, C.c_float, C.c_float, _P]),
    "pt_trace_closest": (C.c_int, [_P, _P, _U32, _P]),
    "pt_trace_any": (C.c_int, [_P, _P, _U32, _P]),
    "pt_trace_closest_device": (C.c_int, [_P, _P, _U32, _P, _P]),
    "pt_trace_any_device": (C.c_int, [_P, _P, _U32, _P, _P]),
    "pt_trace_closest_counts": (C.c_int, [_P, _P, _U32, _P, _P, _P]),
    "pt_trace_any_counts": (C.c_int, [_P, _P, _U32, _P, _P, _P]),
    "pt_trace_closest_device_noalloc": (C.c_int, [_P, _P, _U32, _P, _P]),
    "pt_frame_size": (None, [_P, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
    "pt_get_queue_occupancy": (C.c_int, [_P, _P, _U32, C.POINTER(C.c_uint32)]),
    "pt_classify_hits": (C.c_int, [_P, _P, _U32, _P, _P]),
    "pt_alias_sample": (C.c_int, [_P, _P, _U32, _P, _P]),
    "pt_get_alias_table": (C.c_int, [_P, _P, _U32, C.POINTER(_U32)]),
    "pt_get_bvh_info": (C.c_int, [_P, C.POINTER(BvhInfo)]),
    "pt_get_bvh4": (C.c_int, [_P, _P, _U32, _P, _U32]),
    "pt_build_bvh4": (C.c_int, [_P, _U32, _U32, _P, _U32, C.POINTER(_U32), _P, C.POINTER(_U32)]),
    "pt_import_obj": (C.c_int, [C.c_char_p, C.POINTER(_P), C.c_char_p, C.c_size_t]),
    "pt_imported_scene_desc": (C.POINTER(SceneDesc), [_P]),
    "pt_imported_scene_free": (None, [_P]),
    "pt_generate_tangents": (C.c_int, [_P, _P, _P, C.c_uint32, _P, C.c_uint32, _P]),
    "pt_write_image": (C.c_int, [C.c_char_p, C.c_uint32, C.c_uint32, C.c_uint32, _P, C.c_int]),
    "pt_save_frame": (C.c_int, [_P, C.c_char_p]),
    "pt_quantize_bvh4": (C.c_int, [_P, C.c_uint32, _P]),
    "pt_build_bvh8_quantized": (C.c_int, [_P, C.c_uint32, C.c_uint32, _P, C.c_uint32, C.POINTER(C.c_uint32), _P,
                                          C.POINTER(C.c_uint32)]),
    "pt_get_bvh4_quantized": (C.c_int, [_P, _P, C.c_uint32]),
    "pt_get_bvh8_quantized": (C.c_int, [_P, _P, C.c_uint32, C.POINTER(C.c_uint32), _P]),
    "pt_get_tuning": (C.c_int, [_P, C.POINTER(Tuning)]),
    "pt_set_tuning": (C.c_int, [_P, C.POINTER(Tuning)]),
}

_lib = None


def load_library(path=LIB_PATH):
    """Load libptcore.so and bind every declared entry point."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: build it with `python gpu-path-tracer_amd/build.py`")
    lib = C.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def _settings(d):
    s = RenderSettings()
    s.width = d.get("width", 256)
    s.height = d.get("height", 256)
    s.samples_per_pixel = d.get("samples_per_pixel", 16)
    s.path_depth = d.get("path_depth", 8)
    s.use_nee = d.get("use_nee", 1)
    s.rr_start_depth = d.get("rr_start_depth", 3)
    s.seed = d.get("seed", 1)
    return s


def _motion(view):
    m = MotionState()
    m.position[:] = [float(c) for c in view.position]
    m.rotation[:] = [float(c) for c in view.rotation]
    return m


def _scene_desc(scene):
    """pt_scene_desc of a scenes.Scene, and the arrays it points into."""
    arrs = dict(
        positions=np.ascontiguousarray(scene.positions, np.float32),
        normals=np.ascontiguousarray(scene.normals, np.float32),
        tangents=None if scene.tangents is None else np.ascontiguousarray(scene.tangents, np.float32),
        material_ids=np.ascontiguousarray(scene.material_ids, np.uint32),
        materials=np.ascontiguousarray(scene.materials),
        light_weights=None if scene.light_weights is None else np.ascontiguousarray(scene.light_weights, np.float32),
    )
    d = SceneDesc(scene.triangle_count, _ptr(arrs["positions"]), _ptr(arrs["normals"]), _ptr(arrs["tangents"]),
                  _ptr(arrs["material_ids"]), len(arrs["materials"]), _ptr(arrs["materials"]),
                  _ptr(arrs["light_weights"]))
    return d, arrs


def _view(scene_id, settings, view):
    v = View()
    v.scene = scene_id
    v.settings = _settings(settings)
    v.camera = Camera(view.fov_y, 0.1, 1000.0)
    v.state = _motion(view)
    return v


class Renderer:
    """Python mirror of Core::Graphics::Cuda::Renderer over the C ABI.  Method
    names follow the reference; each failure raises PtError with the status
    and the library's message (the reference returns std::unexpected(Error))."""

    def __init__(self):
        self.lib = load_library()
        h = C.c_void_p()
        st = self.lib.pt_renderer_create(C.byref(h))
        if st != PT_OK:
            raise PtError(st, "pt_renderer_create failed (is a GPU visible?)")
        self.h = h
        self._keep = None

    def close(self):
        if self.h:
            self.lib.pt_renderer_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _check(self, st):
        if st != PT_OK:
            msg = self.lib.pt_get_error_message(self.h)
            raise PtError(st, msg.decode() if msg else "")

    # ---- boundary (Renderer.hpp:114-143)
    def InitializeRenderingBuffers(self, width, height):
        self._check(self.lib.pt_initialize_rendering_buffers(self.h, width, height))

    def InitializeSceneBuffers(self, scene):
        d, _arrs = _scene_desc(scene)
        self._check(self.lib.pt_initialize_scene_buffers(self.h, C.byref(d)))

    def Clear(self, r, g, b, a):
        self._check(self.lib.pt_clear(self.h, r, g, b, a))

    def StartSimulation(self, settings, view, motion_states, output_folder=None):
        s = _settings(settings)
        cam = Camera(view.fov_y, 0.1, 1000.0)
        arr = (MotionState * len(motion_states))(*[_motion(v) for v in motion_states])
        self._keep = arr
        out = output_folder.encode() if output_folder else None
        self._check(self.lib.pt_start_simulation(self.h, C.byref(s), C.byref(cam), arr, len(motion_states), out))

    def ResumeSimulation(self, start_frame):
        self._check(self.lib.pt_resume_simulation(self.h, start_frame))

    def StopRendering(self):
        self._check(self.lib.pt_stop_rendering(self.h))

    def IsRendering(self):
        return bool(self.lib.pt_is_rendering(self.h))

    def GetProgress(self):
        p = Progress()
        self._check(self.lib.pt_get_progress(self.h, C.byref(p)))
        return {k: getattr(p, k) for k, _ in p._fields_}

    def GetFrame(self, width, height):
        out = np.zeros(width * height, np.uint32)
        self._check(self.lib.pt_frame_copy(self.h, _ptr(out), out.size))
        return out.reshape(height, width)

    @contextlib.contextmanager
    def LockFrame(self):
        """pt_frame_lock / pt_frame_unlock: the published RGBA8 frame without a
        copy, as a (pixels,) uint32 view valid only inside the ``with`` block
        (FrameView::Lock(), Renderer.hpp:45-104)."""
        ptr = C.POINTER(C.c_uint32)()
        n = C.c_size_t()
        self._check(self.lib.pt_frame_lock(self.h, C.byref(ptr), C.byref(n)))
        try:
            yield np.ctypeslib.as_array(ptr, (n.value,)) if n.value else np.zeros(0, np.uint32)
        finally:
            self._check(self.lib.pt_frame_unlock(self.h))

    def GetFrameHdr(self, width, height):
        out = np.zeros((height, width, 4), np.float32)
        self._check(self.lib.pt_frame_copy_hdr(self.h, _ptr(out), width * height))
        return out

    def SaveFrame(self, path):
        """pt_save_frame: ".png" (RGBA8 frame) or ".exr" (float radiance)."""
        self._check(self.lib.pt_save_frame(self.h, os.fspath(path).encode()))

    def PeekLastError(self):
        buf = C.create_string_buffer(4096)
        n = self.lib.pt_peek_last_error(self.h, buf, 4096)
        return buf.value.decode() if n else None

    def ConsumeLastError(self):
        buf = C.create_string_buffer(4096)
        n = self.lib.pt_consume_last_error(self.h, buf, 4096)
        return buf.value.decode() if n else None

    # ---- synchronous frame
    def RenderFrame(self, settings, view, flags=FRAME_DEFAULT):
        s = _settings(settings)
        cam = Camera(view.fov_y, 0.1, 1000.0)
        m = _motion(view)
        st = FrameStats()
        self._check(self.lib.pt_render_frame(self.h, C.byref(s), C.byref(cam), C.byref(m), flags, C.byref(st)))
        return st.as_dict()

    # ---- resident scenes and batched views
    def AddScene(self, scene):
        """Uploads a further scene; returns its id (1..63)."""
        d, _arrs = _scene_desc(scene)
        sid = C.c_uint32()
        self._check(self.lib.pt_add_scene(self.h, C.byref(d), C.byref(sid)))
        return sid.value

    def RemoveScene(self, scene_id):
        self._check(self.lib.pt_remove_scene(self.h, scene_id))

    def RenderBatch(self, items, flags=FRAME_DEFAULT, rgba=False, hdr=True):
        """items: [(scene_id, settings dict, scenes.View)], rendered in one
        wavefront.  Returns ([hdr (h, w, 4) per view, or None when hdr=False],
        stats), plus the RGBA8 images (h, w) when rgba=True.  hdr may also be
        a list of booleans, one per view."""
        n = len(items)
        views = (View * n)(*[_view(sid, st, v) for sid, st, v in items])
        want_hdr = hdr if isinstance(hdr, (list, tuple)) else [hdr] * n
        hdr, ldr = [], []
        for i in range(n):
            w, h = self.frame_size(views[i].settings)
            hdr.append(np.zeros((h, w, 4), np.float32) if want_hdr[i] else None)
            ldr.append(np.zeros((h, w), np.uint32) if rgba else None)
        hp = (C.c_void_p * n)(*[a.ctypes.data if a is not None else None for a in hdr])
        lp = (C.c_void_p * n)(*[a.ctypes.data if a is not None else None for a in ldr])
        st = FrameStats()
        self._check(self.lib.pt_render_batch(self.h, views, n, hp, lp, flags, C.byref(st)))
        return (hdr, st.as_dict(), ldr) if rgba else (hdr, st.as_dict())

    # ---- learned path guiding
    def RenderGuided(self, settings, view, passes=4, train_steps=32, batch=16384, learning_rate=1e-3,
                     guide_fraction=0.5, weight_max=1e4, record_capacity=0, reset=True, seed=1):
        """Renders a frame in `passes` passes with the guide trained between
        them; returns (frame stats, guide stats)."""
        s = _settings(settings)
        cam = Camera(view.fov_y, 0.1, 1000.0)
        m = _motion(view)
        g = GuideSettings(passes, train_steps, batch, learning_rate, guide_fraction, weight_max, record_capacity,
                          1 if reset else 0, seed)
        gst, st = GuideStats(), FrameStats()
        self._check(self.lib.pt_render_guided(self.h, C.byref(s), C.byref(cam), C.byref(m), C.byref(g), C.byref(gst),
                                              C.byref(st)))
        return st.as_dict(), {k: getattr(gst, k) for k, _ in gst._fields_}

    def guide_reset(self, seed=1):
        self._check(self.lib.pt_guide_reset(self.h, seed))

    def guide_weights(self):
        w = np.zeros(GUIDE_PARAMS, np.float32)
        self._check(self.lib.pt_guide_get_weights(self.h, _ptr(w), GUIDE_PARAMS))
        return w

    def guide_set_weights(self, w):
        w = np.ascontiguousarray(w, np.float32)
        self._check(self.lib.pt_guide_set_weights(self.h, _ptr(w), len(w)))

    def guide_set_box(self, lo, hi):
        lo = np.ascontiguousarray(lo, np.float32)
        hi = np.ascontiguousarray(hi, np.float32)
        self._check(self.lib.pt_guide_set_box(self.h, _ptr(lo), _ptr(hi)))

    def guide_forward(self, inputs):
        x = np.ascontiguousarray(inputs, np.float32)
        out = np.zeros((len(x), GUIDE_DIM), np.float32)
        self._check(self.lib.pt_guide_forward(self.h, _ptr(x), len(x), _ptr(out)))
        return out

    def guide_mixture(self, inputs):
        """(mu [n][8][3], kappa [n][8], pi [n][8]) as the inference kernel stores them."""
        x = np.ascontiguousarray(inputs, np.float32)
        out = np.zeros((len(x), 8, 5), np.float32)
        self._check(self.lib.pt_guide_mixture(self.h, _ptr(x), len(x), _ptr(out)))
        return out[..., :3], out[..., 3], out[..., 4]

    def guide_gradient(self, records, weight_max=1e30):
        r = np.ascontiguousarray(records, GUIDE_RECORD_DTYPE)
        g = np.zeros(GUIDE_PARAMS, np.float32)
        loss = np.zeros(1, np.float32)
        self._check(self.lib.pt_guide_gradient(self.h, _ptr(r), len(r), weight_max, _ptr(g), _ptr(loss)))
        return g, float(loss[0])

    def guide_train(self, records, steps, batch, learning_rate=1e-3, weight_max=1e30):
        r = np.ascontiguousarray(records, GUIDE_RECORD_DTYPE)
        losses = np.zeros(steps, np.float32)
        self._check(self.lib.pt_guide_train(self.h, _ptr(r), len(r), steps, batch, learning_rate, weight_max,
                                            _ptr(losses)))
        return losses

    # ---- serving
    def ServerStart(self, accum_budget_bytes=0, window_samples=0, chunk_samples=0, max_requests=0):
        cfg = ServerConfig(accum_budget_bytes, window_samples, chunk_samples, max_requests, 0)
        self._check(self.lib.pt_server_start(self.h, C.byref(cfg)))
        self._requests = {}

    def ServerSubmit(self, scene_id, settings, view, hdr=True, rgba=False):
        """Queues one view; returns its request id.  The result arrays are
        held here until ServerResult takes them."""
        v = _view(scene_id, settings, view)
        w, h = self.frame_size(v.settings)
        out_hdr = np.zeros((h, w, 4), np.float32) if hdr else None
        out_rgba = np.zeros((h, w), np.uint32) if rgba else None
        rid = C.c_uint64()
        self._check(self.lib.pt_server_submit(self.h, C.byref(v), _ptr(out_hdr), _ptr(out_rgba), C.byref(rid)))
        self._requests[rid.value] = (out_hdr, out_rgba)
        return rid.value

    def frame_size(self, settings):
        """(width, height) of a frame rendered with `settings` (a RenderSettings),
        as the library clamps them: the size of every result buffer."""
        w, h = C.c_uint32(), C.c_uint32()
        self.lib.pt_frame_size(C.byref(settings), C.byref(w), C.byref(h))
        return w.value, h.value

    def ServerWait(self, request_id, timeout_ms=120000):
        info = RequestInfo()
        self._check(self.lib.pt_server_wait(self.h, request_id, timeout_ms, C.byref(info)))
        return {k: getattr(info, k) for k, _ in info._fields_}

    def ServerResult(self, request_id):
        """(hdr, rgba) arrays of a finished request."""
        return self._requests.pop(request_id)

    def ServerStats(self):
        st = ServerStats()
        self._check(self.lib.pt_server_get_stats(self.h, C.byref(st)))
        return {k: getattr(st, k) for k, _ in st._fields_}

    def ServerStop(self):
        st = ServerStats()
        self._check(self.lib.pt_server_stop(self.h, C.byref(st)))
        return {k: getattr(st, k) for k, _ in st._fields_}

    # ---- kernel-level hooks
    def trace_closest(self, rays):
        rays = np.ascontiguousarray(rays, np.float32)
        hits = np.zeros((len(rays), 4), np.uint32)
        self._check(self.lib.pt_trace_closest(self.h, _ptr(rays), len(rays), _ptr(hits)))
        return hits

    def trace_any(self, rays):
        rays = np.ascontiguousarray(rays, np.float32)
        occ = np.zeros(len(rays), np.uint32)
        self._check(self.lib.pt_trace_any(self.h, _ptr(rays), len(rays), _ptr(occ)))
        return occ

    def trace_closest_counts(self, rays):
        rays = np.ascontiguousarray(rays, np.float32)
        hits = np.zeros((len(rays), 4), np.uint32)
        nodes = np.zeros(len(rays), np.uint32)
        tris = np.zeros(len(rays), np.uint32)
        self._check(self.lib.pt_trace_closest_counts(self.h, _ptr(rays), len(rays), _ptr(hits), _ptr(nodes),
                                                     _ptr(tris)))
        return hits, nodes, tris

    def trace_any_counts(self, rays):
        rays = np.ascontiguousarray(rays, np.float32)
        occ = np.zeros(len(rays), np.uint32)
        nodes = np.zeros(len(rays), np.uint32)
        tris = np.zeros(len(rays), np.uint32)
        self._check(self.lib.pt_trace_any_counts(self.h, _ptr(rays), len(rays), _ptr(occ), _ptr(nodes), _ptr(tris)))
        return occ, nodes, tris

    def classify(self, rays):
        rays = np.ascontiguousarray(rays, np.float32)
        q = np.zeros(len(rays), np.uint32)
        counts = np.zeros(5, np.uint32)
        self._check(self.lib.pt_classify_hits(self.h, _ptr(rays), len(rays), _ptr(q), _ptr(counts)))
        return q, counts

    def alias_sample(self, u):
        u = np.ascontiguousarray(u, np.float32)
        idx = np.zeros(len(u), np.uint32)
        pmf = np.zeros(len(u), np.float32)
        self._check(self.lib.pt_alias_sample(self.h, _ptr(u), len(u), _ptr(idx), _ptr(pmf)))
        return idx, pmf

    def alias_table(self):
        n = C.c_uint32()
        self._check(self.lib.pt_get_alias_table(self.h, None, 0, C.byref(n)))
        bins = np.zeros(n.value, ALIAS_DTYPE)
        if n.value:
            self._check(self.lib.pt_get_alias_table(self.h, _ptr(bins), n.value, C.byref(n)))
        return bins

    def bvh_info(self):
        b = BvhInfo()
        self._check(self.lib.pt_get_bvh_info(self.h, C.byref(b)))
        return {k: getattr(b, k) for k, _ in b._fields_}

    def queue_occupancy(self):
        """Per-iteration queue sizes of the last FRAME_PROFILE frame
        (QUEUE_OCCUPANCY_DTYPE records)."""
        n = C.c_uint32(0)
        self._check(self.lib.pt_get_queue_occupancy(self.h, None, 0, C.byref(n)))
        out = np.zeros(n.value, QUEUE_OCCUPANCY_DTYPE)
        if n.value:
            self._check(self.lib.pt_get_queue_occupancy(self.h, _ptr(out), n.value, C.byref(n)))
        return out

    def bvh4(self):
        info = self.bvh_info()
        nodes = np.zeros((info["node_count"], 32), np.uint32)
        order = np.zeros(info["triangle_count"], np.uint32)
        self._check(self.lib.pt_get_bvh4(self.h, _ptr(nodes), info["node_count"], _ptr(order), len(order)))
        return nodes, order

    def bvh8_quantized(self):
        """(N, 32) uint32 rows of the 128-byte 8-wide quantized nodes of scene 0 and
        the tree's triangle order (original prim of each position)."""
        n = C.c_uint32(0)
        self._check(self.lib.pt_get_bvh8_quantized(self.h, None, 0, C.byref(n), None))
        out = np.zeros((n.value, 32), np.uint32)
        order = np.zeros(self.bvh_info()["triangle_count"], np.uint32)
        self._check(self.lib.pt_get_bvh8_quantized(self.h, _ptr(out), n.value, C.byref(n), _ptr(order)))
        return out, order

    def bvh4_quantized(self):
        """(N, 16) uint32 rows of the 64-byte quantized nodes scene 0 is traversed with."""
        n = self.bvh_info()["node_count"]
        out = np.zeros((n, 16), np.uint32)
        self._check(self.lib.pt_get_bvh4_quantized(self.h, _ptr(out), n))
        return out

    def get_tuning(self):
        t = Tuning()
        self._check(self.lib.pt_get_tuning(self.h, C.byref(t)))
        return {k: getattr(t, k) for k, _ in t._fields_}

    def set_tuning(self, **kw):
        cur = self.get_tuning()
        cur.update(kw)
        t = Tuning(**cur)
        self._check(self.lib.pt_set_tuning(self.h, C.byref(t)))


def build_bvh4(positions, max_leaf_size=0):
    """Host-only BVH4 build: returns (nodes (N, 32) uint32 rows of 128 B, prim order, depth)."""
    lib = load_library()
    P = np.ascontiguousarray(positions, np.float32)
    n, depth = C.c_uint32(), C.c_uint32()
    st = lib.pt_build_bvh4(_ptr(P), len(P), max_leaf_size, None, 0, C.byref(n), None, C.byref(depth))
    if st != PT_OK:
        raise PtError(st, "pt_build_bvh4 failed")
    nodes = np.zeros((n.value, 32), np.uint32)
    order = np.zeros(len(P), np.uint32)
    st = lib.pt_build_bvh4(_ptr(P), len(P), max_leaf_size, _ptr(nodes), n.value, C.byref(n), _ptr(order),
                           C.byref(depth))
    if st != PT_OK:
        raise PtError(st, "pt_build_bvh4 failed")
    return nodes, order, depth.value


def build_bvh8_quantized(positions, max_leaf_size=0):
    """Host-only 8-wide quantized build: (nodes (N, 32) uint32 rows of 128 B,
    the tree's triangle order (original prim of each position), depth)."""
    lib = load_library()
    P = np.ascontiguousarray(positions, np.float32)
    n, depth = C.c_uint32(), C.c_uint32()
    st = lib.pt_build_bvh8_quantized(_ptr(P), len(P), max_leaf_size, None, 0, C.byref(n), None, C.byref(depth))
    if st != PT_OK:
        raise PtError(st, "pt_build_bvh8_quantized failed")
    nodes = np.zeros((n.value, 32), np.uint32)
    order = np.zeros(len(P), np.uint32)
    st = lib.pt_build_bvh8_quantized(_ptr(P), len(P), max_leaf_size, _ptr(nodes), n.value, C.byref(n), _ptr(order),
                                     C.byref(depth))
    if st != PT_OK:
        raise PtError(st, "pt_build_bvh8_quantized failed")
    return nodes, order, depth.value


def quantize_bvh4(nodes):
    """pt_quantize_bvh4: (N, 32) uint32 rows of 128-byte BVH4 nodes -> (N, 16)
    uint32 rows of 64-byte quantized nodes."""
    lib = load_library()
    a = np.ascontiguousarray(nodes, np.uint32).reshape(-1, 32)
    out = np.zeros((len(a), 16), np.uint32)
    st = lib.pt_quantize_bvh4(_ptr(a), len(a), _ptr(out))
    if st != PT_OK:
        raise PtError(st, "pt_quantize_bvh4 failed")
    return out


def import_obj(path):
    """pt_import_obj: an OBJ file and its MTL libraries as a scenes.Scene
    (arrays copied out of the library's import)."""
    import scenes
    lib = load_library()
    h = _P()
    err = C.create_string_buffer(1024)
    st = lib.pt_import_obj(os.fspath(path).encode(), C.byref(h), err, len(err))
    if st != PT_OK:
        raise PtError(st, err.value.decode(errors="replace"))
    try:
        d = lib.pt_imported_scene_desc(h).contents
        T, M = d.triangle_count, d.material_count

        def arr(ptr, dtype, shape):
            n = int(np.prod(shape))
            if n == 0:
                return np.zeros(shape, dtype)
            buf = (C.c_char * (n * np.dtype(dtype).itemsize)).from_address(ptr)
            return np.frombuffer(buf, dtype).reshape(shape).copy()

        return scenes.Scene(name=os.path.basename(os.fspath(path)), positions=arr(d.positions, np.float32, (T, 3, 3)),
                            normals=arr(d.normals, np.float32, (T, 3, 3)),
                            tangents=arr(d.tangents, np.float32, (T, 3, 4)),
                            material_ids=arr(d.material_ids, np.uint32, (T,)),
                            materials=arr(d.materials, scenes.MATERIAL_DTYPE, (M,)))
    finally:
        lib.pt_imported_scene_free(h)


def generate_tangents(positions, normals, uvs, indices):
    """pt_generate_tangents: (V, 4) float32 per-vertex tangents (xyz, sign) of
    the triangles `indices` over per-vertex positions, unit normals and uvs."""
    lib = load_library()
    P = np.ascontiguousarray(positions, np.float32).reshape(-1, 3)
    N = np.ascontiguousarray(normals, np.float32).reshape(-1, 3)
    U = np.ascontiguousarray(uvs, np.float32).reshape(-1, 2)
    idx = np.ascontiguousarray(indices, np.uint32).reshape(-1)
    if not (len(P) == len(N) == len(U)):
        raise ValueError("positions, normals and uvs need one row per vertex")
    out = np.zeros((len(P), 4), np.float32)
    st = lib.pt_generate_tangents(_ptr(P), _ptr(N), _ptr(U), len(P), _ptr(idx), len(idx), _ptr(out))
    if st != PT_OK:
        raise PtError(st, "pt_generate_tangents: index count not a multiple of 3 or an index out of range")
    return out


PT_PIXELS_RGBA8 = 0
PT_PIXELS_RGBA32F = 1


def write_image(path, pixels, bottom_up=False):
    """pt_write_image: (H, W) uint32 RGBA8 words to ".png", or (H, W, 4)
    float32 RGBA to ".exr"; row 0 is the top unless `bottom_up`."""
    lib = load_library()
    a = np.ascontiguousarray(pixels)
    if a.dtype == np.uint32 and a.ndim == 2:
        fmt, (h, w) = PT_PIXELS_RGBA8, a.shape
    elif a.dtype == np.float32 and a.ndim == 3 and a.shape[2] == 4:
        fmt, (h, w) = PT_PIXELS_RGBA32F, a.shape[:2]
    else:
        raise ValueError("pixels: (H, W) uint32 RGBA8 or (H, W, 4) float32")
    st = lib.pt_write_image(os.fspath(path).encode(), w, h, fmt, _ptr(a), int(bool(bottom_up)))
    if st != PT_OK:
        raise PtError(st, f"pt_write_image({os.fspath(path)!r}) failed")


def hits_as_fields(hits):
    """(n, 4) uint32 pt_hit rows -> prim (u32), t, u, v (f32)."""
    h = np.ascontiguousarray(hits, np.uint32)
    return h[:, 0].copy(), h[:, 1].view(np.float32).copy(), h[:, 2].view(np.float32).copy(), h[:, 3].view(np.float32).copy()
