"""Writes the golden fixtures of tests/golden from the CPU oracle.

    python tests/golden/make_golden.py

c1_hits_4096.npz: 4,096 seeded rays into the 36-triangle box room, their
closest hits (prim, t, u, v) under the (t, prim id) order, queue membership,
and 4,096 shadow segments with their occlusion.  The reference ships no test
vectors, so these pin the oracle (and the GPU) against later changes.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "gpu-path-tracer_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import oracle  # noqa: E402
import scenes  # noqa: E402


def main():
    s = scenes.cornell_box()
    o = oracle.OracleScene(s)
    rays = scenes.random_rays(4096, (-1, 0, -1), (1, 2, 1), 4096)
    sh = scenes.shadow_rays(s, 4096, 4097)
    np.savez_compressed(os.path.join(HERE, "c1_hits_4096.npz"), rays=rays, hits=o.closest(rays, brute=True),
                        queue=o.classify(rays), shadow_rays=sh, occluded=o.any(sh))
    w = np.random.default_rng(1024).lognormal(0, 1, 1024).astype(np.float32)
    bins = oracle.alias_build(w)
    u = np.random.default_rng(1025).random(4096).astype(np.float32)
    idx, pmf = oracle.alias_sample(bins, u)
    np.savez_compressed(os.path.join(HERE, "alias_1024.npz"), weights=w, bins=bins.view(np.uint32).reshape(-1, 4),
                        u=u, index=idx, pmf=pmf)
    st = dict(width=16, height=16, samples_per_pixel=64, path_depth=8, use_nee=1, rr_start_depth=3, seed=1)
    hdr, ext, shc = o.render(st, s.view)
    np.savez_compressed(os.path.join(HERE, "c1_render_16x16_64spp.npz"), hdr=hdr, ext=ext, shadow=shc)
    print("wrote fixtures to", HERE)


if __name__ == "__main__":
    main()
