"""The host-side frame of the single-frame API (Renderer.hpp:114-143:
Clear, RenderFrame's copy of the finished frame to the host, the frame copy a
caller reads): every finished render becomes the frame a reader copies, a
size change starts from a cleared frame of the new size, Clear paints the
current frame, and a caller buffer smaller than the frame is refused."""
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
