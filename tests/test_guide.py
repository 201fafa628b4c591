"""Learned path guiding on the GPU: the BF16 MFMA forward and backward passes
against the CPU oracle, online training on a known distribution, and the
guided renderer (unbiased against plain BSDF sampling, lower error at equal
samples)."""
import math
import os
import sys

import numpy as np
import pytest

import ptcore
import scenes

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import guide_oracle as G  # noqa: E402

pytestmark = pytest.mark.gpu


def _unit(v):
    return v / np.linalg.norm(v, axis=1, keepdims=True)


def _records(n, seed):
    rng = np.random.default_rng(seed)
    r = np.zeros(n, ptcore.GUIDE_RECORD_DTYPE)
    r["position"] = rng.uniform(0.0, 1.0, (n, 3))
    r["normal"] = _unit(rng.normal(size=(n, 3)))
    r["wo"] = _unit(rng.normal(size=(n, 3)))
    r["wi"] = _unit(rng.normal(size=(n, 3)))
    r["radiance"] = rng.uniform(0.0, 2.0, (n, 3))
    r["pdf"] = rng.uniform(0.05, 1.0, n)
    r["cos_i"] = rng.uniform(0.0, 1.0, n)
    r["prev"] = 0xFFFFFFFF
    return r


def test_guide_forward_matches_oracle(gpu_renderer):
    w = G.init_weights(1)
    gpu_renderer.guide_set_weights(w)
    assert np.array_equal(gpu_renderer.guide_weights(), w)
    x = np.random.default_rng(2).normal(0.0, 1.0, (300, 64)).astype(np.float32)   # 300: a partial last tile
    got = gpu_renderer.guide_forward(x)
    want, _ = G.forward(w, x, round=True)
    err = np.abs(got - want).max() / np.abs(want).max()
    assert err < 1e-2, err
    # rows are independent samples: permuting the inputs permutes the outputs
    perm = np.random.default_rng(3).permutation(300)
    assert np.allclose(gpu_renderer.guide_forward(x[perm]), got[perm], rtol=1e-6, atol=1e-6)


def test_guide_mixture_matches_oracle(gpu_renderer):
    """The inference kernel's stored mixtures (fp16) against the oracle's
    mixture transform of the same network: the renderer's kernel on inputs
    given as features, compared with the oracle's bf16 forward pass and with
    the mixture of the MFMA forward's own raw outputs (which isolates the
    output transform and the fp16 packing)."""
    w = G.init_weights(6, scale=3.0)
    gpu_renderer.guide_set_weights(w)
    rng = np.random.default_rng(9)
    n = 1000                                                   # a partial last tile
    x = G.features(np.zeros(3), np.ones(3), rng.uniform(0.0, 1.0, (n, 3)), _unit(rng.normal(size=(n, 3))),
                   _unit(rng.normal(size=(n, 3)))).astype(np.float32)
    mu, kappa, pi = gpu_renderer.guide_mixture(x)
    mu_r, kappa_r, pi_r = G.mixture(gpu_renderer.guide_forward(x))
    assert np.abs(mu - mu_r).max() < 1e-3
    assert np.abs(kappa / kappa_r - 1.0).max() < 1e-3
    assert np.abs(pi - pi_r).max() < 1e-3
    assert np.abs(np.linalg.norm(mu, axis=-1) - 1.0).max() < 2e-3
    assert np.abs(pi.sum(-1) - 1.0).max() < 4e-3
    mu_o, kappa_o, pi_o = G.mixture(G.forward(w, x, round=True)[0])
    assert kappa_r.std() > 0.1 and pi_r.std() > 0.01           # the transform sees spread-out outputs
    assert np.abs(pi - pi_o).max() < 2e-2
    assert np.median(np.abs(mu - mu_o)) < 1e-2
    assert np.median(np.abs(kappa / kappa_o - 1.0)) < 2e-2


def test_guide_gradient_matches_oracle(gpu_renderer):
    w = G.init_weights(4)
    gpu_renderer.guide_set_weights(w)
    gpu_renderer.guide_set_box([0.0, 0.0, 0.0], [1.0, 1.0, 1.0])
    recs = _records(640, 5)
    g, loss = gpu_renderer.guide_gradient(recs)
    gref, lref = G.record_gradient(w, recs, np.zeros(3), np.ones(3), round=True)
    assert abs(loss - lref) <= 2e-3 * abs(lref), (loss, lref)
    for l in range(G.LAYERS):
        base = l * G.LAYER_PARAMS
        for part in (slice(base, base + 4096), slice(base + 4096, base + G.LAYER_PARAMS)):
            rel = np.linalg.norm(g[part] - gref[part]) / np.linalg.norm(gref[part])
            assert rel < 3e-2, (l, part, rel)
    assert np.array_equal(gpu_renderer.guide_weights(), w)      # the gradient leaves the weights alone


def test_guide_learns_a_directional_distribution(gpu_renderer):
    """Incident directions drawn from vMF(mu(p), 20) with mu depending on the
    position: training must raise the held-out log-likelihood far above the
    uniform density's log(1/4pi)."""
    rng = np.random.default_rng(8)

    def make(n):
        r = np.zeros(n, ptcore.GUIDE_RECORD_DTYPE)
        p = rng.uniform(0.0, 1.0, (n, 3))
        mu = _unit(np.stack([p[:, 0] - 0.5, np.ones(n), p[:, 2] - 0.5], 1))
        r["position"] = p
        r["normal"] = [0.0, 1.0, 0.0]
        r["wo"] = [0.0, 1.0, 0.0]
        r["wi"] = G.vmf_sample(mu, 20.0, rng.random(n), rng.random(n))
        r["radiance"] = 1.0
        r["pdf"] = 1.0 / (4.0 * np.pi)
        r["cos_i"] = 1.0
        return r

    train, held = make(1 << 16), make(4096)
    gpu_renderer.guide_reset(3)
    gpu_renderer.guide_set_box([0.0, 0.0, 0.0], [1.0, 1.0, 1.0])
    losses = gpu_renderer.guide_train(train, steps=400, batch=4096, learning_rate=3e-3)
    assert losses[-20:].mean() < losses[:20].mean()
    x = G.features(np.zeros(3), np.ones(3), held["position"], held["normal"], held["wo"]).astype(np.float32)
    lq = G.log_q(gpu_renderer.guide_forward(x), held["wi"])
    assert lq.mean() > math.log(1.0 / (4.0 * math.pi)) + 2.0, lq.mean()


@pytest.fixture(scope="module")
def room_renderer():
    room = scenes.guiding_room(20_000)
    r = ptcore.Renderer()
    r.InitializeRenderingBuffers(320, 180)
    r.InitializeSceneBuffers(room)
    yield r, room
    r.close()


def _nmse(a, ref):
    """Mean squared error over the squared mean of the reference (the room is
    dark, so a per-pixel relative error would be dominated by its regulariser)."""
    return float(np.mean((a - ref) ** 2)) / float(ref.mean()) ** 2


def _nmse_trimmed(a, ref, q=0.999):
    """_nmse with each pixel's squared error capped at its 99.9th percentile:
    at 256 spp without NEE a handful of paths that reach the light through the
    gap carry most of a dark image's squared error, for either sampler, and
    one of them can swing the ratio of two such errors several-fold."""
    e = ((a - ref) ** 2).mean(-1).ravel()
    return float(np.minimum(e, np.quantile(e, q)).mean()) / float(ref.mean()) ** 2


def _render_set(r, room, w, h, nee):
    """Reference (16,384 spp), plain and guided (online-trained, then the trained
    guide alone) images at 256 spp."""
    base = dict(width=w, height=h, path_depth=8, use_nee=nee, rr_start_depth=3)
    r.RenderGuided(dict(base, samples_per_pixel=16384, seed=99), room.view, passes=4, guide_fraction=0.0)
    ref = r.GetFrameHdr(w, h)[..., :3].copy()
    st = dict(base, samples_per_pixel=256, seed=7)
    sp, gp = r.RenderGuided(st, room.view, passes=4, guide_fraction=0.0)
    plain = r.GetFrameHdr(w, h)[..., :3].copy()
    assert gp["shaded_entries"] == 0        # no guided pass: nothing inferred
    sg, gs = r.RenderGuided(st, room.view, passes=4, train_steps=512, batch=16384, learning_rate=3e-3,
                            guide_fraction=0.5, reset=True, seed=1)
    online = r.GetFrameHdr(w, h)[..., :3].copy()
    assert sp["samples"] == sg["samples"] == w * h * 256
    assert gs["steps"] == 3 * 512 and gs["records"] > 0 and math.isfinite(gs["loss_last"])
    # the guided passes (all but the untrained first) shade diffuse or GGX entries,
    # at most one per extension ray traced
    assert 0 < gs["shaded_entries"] <= sg["extension_rays"]
    r.RenderGuided(st, room.view, passes=1, train_steps=0, guide_fraction=0.8, reset=False)
    trained = r.GetFrameHdr(w, h)[..., :3].copy()
    for img in (online, trained):
        assert np.isfinite(img).all() and (img >= 0).all()
    return ref, plain, online, trained


def test_guided_render_is_unbiased(room_renderer):
    """Without NEE nothing is clamped, so every estimator has the same mean:
    the image means agree with the reference's within 3 standard errors of the
    Monte Carlo noise (the spread of per-pixel differences).  With NEE the
    reference clamps shadow-ray radiance to length 20 (Intersect.cu:125-134),
    a bias that guided paths, with steadier throughput, meet less often."""
    r, room = room_renderer
    ref, plain, online, trained = _render_set(r, room, 320, 180, nee=0)
    for name, img in (("plain", plain), ("online", online), ("trained", trained)):
        d = (img - ref).mean(-1).ravel()
        assert abs(d.mean()) < 3 * d.std() / math.sqrt(d.size), (name, d.mean(), d.std())


@pytest.mark.parametrize("nee, trained_max, online_max", [(1, 0.75, 0.85), (0, 0.4, 0.5)])
def test_trained_guide_gains_over_bsdf_sampling(room_renderer, nee, trained_max, online_max):
    """What the network gains over BSDF sampling at equal samples: the error of
    the online-trained render and of the trained guide alone, relative to
    plain BSDF sampling, with the error of the brightest 0.1 % of pixels capped
    (_nmse_trimmed; measured over five runs: 0.54-0.57 and 0.57-0.62 with NEE,
    0.23-0.31 and 0.11-0.15 without; untrimmed 0.55-0.57 and 0.56-0.62, 0.26-0.35
    and 0.14-0.19)."""
    r, room = room_renderer
    ref, plain, online, trained = _render_set(r, room, 320, 180, nee=nee)
    e_plain, e_online, e_trained = (_nmse_trimmed(plain, ref), _nmse_trimmed(online, ref),
                                    _nmse_trimmed(trained, ref))
    assert e_trained < trained_max * e_plain, (e_trained, e_plain)
    assert e_online < online_max * e_plain, (e_online, e_plain)
