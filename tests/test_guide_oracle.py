"""Pins the guide oracle (oracle/guide_oracle.py) by mathematics: vMF
normalisation and mean cosine, finite differences of the loss and of the whole
network, Adam's first step, and bf16 rounding."""
import math
import os
import sys

import numpy as np

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import guide_oracle as G  # noqa: E402


def test_bf16_rounding_closed_forms():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, 1.0 + 2 ** -7 + 2 ** -8, -2.5, 0.0], np.float32)
    want = np.array([1.0, 1.0, 1.0 + 2 ** -7, 1.0 + 2 ** -6, -2.5, 0.0], np.float32)
    # 1+2^-8 is a tie -> even (1.0); 1+3*2^-9 rounds up; 1+2^-7+2^-8 is a tie -> even (1+2^-6)
    assert np.array_equal(G.bf16(x), want)


def test_vmf_density_integrates_to_one():
    c = np.linspace(-1.0, 1.0, 400001)
    for kappa in (0.14, 1.0, 10.0, 1000.0):
        pdf = np.exp(G.vmf_log_pdf(kappa, c))
        total = 2.0 * np.pi * np.trapezoid(pdf, c)
        assert abs(total - 1.0) < 2e-4, (kappa, total)


def test_vmf_sampling_mean_cosine_is_coth_minus_inverse():
    rng = np.random.default_rng(3)
    for kappa in (0.5, 4.0, 60.0):
        mu = rng.normal(size=3)
        mu /= np.linalg.norm(mu)
        n = 200_000
        d = G.vmf_sample(np.tile(mu, (n, 1)), kappa, rng.random(n), rng.random(n))
        cos = d @ mu
        want = 1.0 / math.tanh(kappa) - 1.0 / kappa
        assert abs(cos.mean() - want) < 4 * cos.std() / math.sqrt(n), (kappa, cos.mean(), want)
        # azimuth is uniform around mu: the mean direction is parallel to mu
        perp = d.mean(0) - (d.mean(0) @ mu) * mu
        assert np.linalg.norm(perp) < 5.0 / math.sqrt(n)


def test_loss_gradient_matches_finite_differences():
    rng = np.random.default_rng(7)
    o = rng.normal(0.0, 1.0, (5, G.OUT))
    o[:, 3::5] = rng.uniform(-1.5, 5.0, (5, G.LOBES))     # inside the kappa clamp
    wi = rng.normal(size=(5, 3))
    wi /= np.linalg.norm(wi, axis=1, keepdims=True)
    w = rng.uniform(0.1, 3.0, 5)
    g, _ = G.loss_grad(o, wi, w)
    h = 1e-6
    for k in range(G.OUT):
        op, om = o.copy(), o.copy()
        op[:, k] += h
        om[:, k] -= h
        fd = (-w * G.log_q(op, wi) + w * G.log_q(om, wi)) / (2 * h)
        assert np.allclose(g[:, k], fd, rtol=1e-5, atol=1e-7), k


def test_network_gradient_matches_finite_differences():
    rng = np.random.default_rng(11)
    w = G.init_weights(5).astype(np.float64)
    n = 6
    recs = np.zeros(n, G.RECORD_DTYPE)
    recs["position"] = rng.uniform(0, 1, (n, 3))
    nn = rng.normal(size=(n, 3))
    recs["normal"] = nn / np.linalg.norm(nn, axis=1, keepdims=True)
    wo = rng.normal(size=(n, 3))
    recs["wo"] = wo / np.linalg.norm(wo, axis=1, keepdims=True)
    wi = rng.normal(size=(n, 3))
    recs["wi"] = wi / np.linalg.norm(wi, axis=1, keepdims=True)
    recs["radiance"] = rng.uniform(0.1, 2.0, (n, 3))
    recs["pdf"] = rng.uniform(0.1, 1.0, n)
    recs["cos_i"] = rng.uniform(0.1, 1.0, n)
    lo, inv = np.zeros(3), np.ones(3)
    g, loss = G.record_gradient(w, recs, lo, inv, round=False)

    def f(wv):
        return G.record_gradient(wv, recs, lo, inv, round=False)[1]

    assert math.isclose(loss, f(w), rel_tol=1e-12)
    h = 1e-6
    picks = [0, 17, 64 * 3 + 5, G.DIM * G.DIM + 3,                     # layer 0 weights, bias
             G.LAYER_PARAMS + 100, 2 * G.LAYER_PARAMS + 4000,           # hidden layers
             3 * G.LAYER_PARAMS + 5 * 2 + 1, 3 * G.LAYER_PARAMS + 5 * 4 + 3, 3 * G.LAYER_PARAMS + G.DIM * G.DIM + 9]
    for k in picks:
        wp, wm = w.copy(), w.copy()
        wp[k] += h
        wm[k] -= h
        fd = (f(wp) - f(wm)) / (2 * h)
        assert math.isclose(g[k], fd, rel_tol=1e-4, abs_tol=1e-7), (k, g[k], fd)


def test_adam_first_step_is_a_signed_step():
    rng = np.random.default_rng(2)
    w0 = rng.normal(size=100)
    g = rng.normal(size=100)
    w1, m, v = G.adam_step(w0, g, np.zeros(100), np.zeros(100), 1, 0.01)
    # bias-corrected first step: m_hat = g, v_hat = g^2 -> w0 - lr g / (|g| + eps)
    assert np.allclose(w1, w0 - 0.01 * g / (np.abs(g) + 1e-8), rtol=1e-12, atol=1e-15)


def test_features_layout():
    p = np.array([[0.25, 0.5, 1.0]])
    x = G.features(np.zeros(3), np.ones(3), p, [[0, 0, 1]], [[1, 0, 0]])
    assert x.shape == (1, 64) and np.allclose(x[0, :3], p[0])
    assert np.allclose(x[0, 3:5], [math.sin(math.pi * 0.25), math.cos(math.pi * 0.25)])
    assert np.allclose(x[0, 39:45], [0, 0, 1, 1, 0, 0]) and (x[0, 45:] == 0).all()
