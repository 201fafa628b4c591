"""The host-side frame of the single-frame API (Renderer.hpp:114-143:
Clear, RenderFrame's copy of the finished frame to the host, the frame copy a
caller reads): every finished render becomes the frame a reader copies, a
size change starts from a cleared frame of the new size, Clear paints the
current frame, and a caller buffer smaller than the frame is refused."""
import ctypes

import numpy as np
import pytest

import ptcore
import scenes

pytestmark = pytest.mark.gpu


def _settings(w, h, seed):
    return dict(width=w, height=h, samples_per_pixel=4, path_depth=4, use_nee=1, rr_start_depth=3, seed=seed)


def test_frames_follow_renders_sizes_and_clear(oracle_lib):
    s = scenes.cornell_box()
    r = ptcore.Renderer()
    try:
        r.InitializeRenderingBuffers(32, 32)
        r.InitializeSceneBuffers(s)
        assert (r.GetFrame(32, 32) == 0).all()
        r.Clear(10, 20, 30, 40)
        assert (r.GetFrame(32, 32) == 0x281E140A).all()

        frames = []
        for seed in (1, 2, 3):
            r.RenderFrame(_settings(32, 32, seed), s.view)
            f = r.GetFrame(32, 32)
            hdr = r.GetFrameHdr(32, 32).reshape(-1, 4)
            # the frame read back is this render's tonemap (Image.cu:34-54)
            want = oracle_lib.tonemap(hdr).reshape(32, 32)
            a = np.stack([(f >> (8 * k)) & 0xFF for k in range(4)], -1).astype(int)
            b = np.stack([(want >> (8 * k)) & 0xFF for k in range(4)], -1).astype(int)
            assert np.abs(a - b).max() <= 1 and (a[..., 3] == 255).all()
            frames.append(f)
        assert not np.array_equal(frames[0], frames[1]) and not np.array_equal(frames[1], frames[2])

        # a render of another size: the frame takes that size
        r.RenderFrame(_settings(48, 40, 4), s.view)
        big = r.GetFrame(48, 40)
        assert (big >> 24 == 255).all()
        with pytest.raises(ptcore.PtError) as e:
            r.GetFrame(32, 32)
        assert e.value.status == 1   # PT_ERROR_INVALID_ARGUMENT
        r.Clear(1, 2, 3, 4)
        assert (r.GetFrame(48, 40) == 0x04030201).all()
        r.RenderFrame(_settings(48, 40, 4), s.view)
        assert np.array_equal(r.GetFrame(48, 40), big)

        # back to the first size: a cleared frame until the next render
        r.InitializeRenderingBuffers(32, 32)
        assert (r.GetFrame(32, 32) == 0).all()
        r.RenderFrame(_settings(32, 32, 3), s.view)
        assert np.array_equal(r.GetFrame(32, 32), frames[2])
    finally:
        r.close()


def test_frame_lock_is_the_published_frame_without_a_copy():
    """pt_frame_lock hands out the published frame itself (FrameView::Lock(),
    Renderer.hpp:45-104): it equals the copy, a second lock or a frame-taking
    call on the locking thread is refused with PT_ERROR_INVALID_STATE, and a
    reader on another thread waits until the lock is released."""
    import threading
    import time

    s = scenes.cornell_box()
    r = ptcore.Renderer()
    try:
        r.InitializeRenderingBuffers(32, 32)
        r.InitializeSceneBuffers(s)
        r.RenderFrame(_settings(32, 32, 5), s.view)
        copy = r.GetFrame(32, 32)
        with r.LockFrame() as px:
            assert px.shape == (32 * 32,)
            assert np.array_equal(px.reshape(32, 32), copy)
            for call in (lambda: r.GetFrame(32, 32), lambda: r.Clear(0, 0, 0, 0),
                         lambda: r.RenderFrame(_settings(32, 32, 6), s.view),
                         lambda: r.InitializeRenderingBuffers(16, 16), r.StopRendering):
                with pytest.raises(ptcore.PtError) as e:
                    call()
                assert e.value.status == 2   # PT_ERROR_INVALID_STATE
            with pytest.raises(ptcore.PtError) as e:
                with r.LockFrame():
                    pass
            assert e.value.status == 2

            got = {}
            t = threading.Thread(target=lambda: got.setdefault("f", r.GetFrame(32, 32)))
            t.start()
            time.sleep(0.3)
            assert t.is_alive() and "f" not in got   # blocked on the frame lock
        t.join(30)
        assert not t.is_alive() and np.array_equal(got["f"], copy)
        with pytest.raises(ptcore.PtError) as e:
            r._check(r.lib.pt_frame_unlock(r.h))   # not locked by this thread
        assert e.value.status == 2

        # after the lock: the calls work again and the frame follows the next render
        r.RenderFrame(_settings(32, 32, 6), s.view)
        with r.LockFrame() as px:
            locked = px.reshape(32, 32).copy()
        assert np.array_equal(locked, r.GetFrame(32, 32)) and not np.array_equal(locked, copy)

        # destroying the renderer on the locking thread releases the lock
        ptr = ctypes.POINTER(ctypes.c_uint32)()
        n = ctypes.c_size_t()
        assert r.lib.pt_frame_lock(r.h, ctypes.byref(ptr), ctypes.byref(n)) == 0 and n.value == 32 * 32
    finally:
        r.close()
