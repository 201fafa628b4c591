"""The bounce loop as a graph runs each bounce's any-hit traversal on a side
stream, overlapping the next bounce's closest-hit traversal: the frame must be
the one the sequential loop renders (same ray and sample counts, images equal
up to the order of float atomics).  The production scene spills traversal
stacks past their LDS part, so both traversals use their global spill areas at
the same time."""
import numpy as np
import pytest

import ptcore
import scenes

pytestmark = pytest.mark.gpu


def test_overlapped_loop_equals_sequential_loop(gpu_renderer):
    s = scenes.production_scene()
    st = dict(s.settings, width=960, height=540, samples_per_pixel=16)
    r = gpu_renderer
    r.InitializeRenderingBuffers(960, 540)
    r.InitializeSceneBuffers(s)
    assert r.bvh_info()["stack_total"] > r.bvh_info()["stack_lds"]
    seq = r.RenderFrame(st, s.view, ptcore.FRAME_NO_GRAPH)
    img_seq = r.GetFrameHdr(960, 540)
    for _ in range(3):
        ov = r.RenderFrame(st, s.view)
        for k in ("extension_rays", "shadow_rays", "samples"):
            assert ov[k] == seq[k], (k, ov[k], seq[k])
        img = r.GetFrameHdr(960, 540)
        assert np.allclose(img, img_seq, rtol=1e-4, atol=1e-5)
