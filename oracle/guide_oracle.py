"""CPU oracle of the guide network (test infrastructure only: imported by
tests/ and nothing else).

The reference has no path guiding, so nothing here restates reference code;
each function restates a published method and is pinned by mathematics in
tests/test_guide_oracle.py:
  * vMF density, its stable sampling (W. Jakob 2012, "Numerically stable
    sampling of the von Mises Fisher distribution on S^2") - pinned by the
    normalisation integral and the closed-form mean cosine coth(k) - 1/k;
  * the mixture loss -w log q and its gradient (Mueller et al. 2019, "Neural
    Importance Sampling"; Dong et al. 2023, "Neural Parametric Mixtures") -
    pinned by central finite differences in float64;
  * back-propagation through the 4-layer ReLU MLP - pinned by finite
    differences of the whole network;
  * Adam (Kingma and Ba 2015) - pinned by its closed form on one step.

With round=True the forward and backward passes round exactly where the
GPU kernels do (bf16 weights, bf16 layer inputs and bf16 output deltas, fp32
or wider accumulation), so GPU results can be compared at tight tolerance.
"""
import numpy as np

DIM = 64
LAYERS = 4
LOBES = 8
OUT = 5 * LOBES
OCTAVES = 6
FEATURES = 3 + 3 * 2 * OCTAVES + 3 + 3
LAYER_PARAMS = DIM * DIM + DIM
PARAMS = LAYERS * LAYER_PARAMS
LOG_KAPPA_MIN, LOG_KAPPA_MAX = -2.0, 7.0

RECORD_DTYPE = np.dtype([("position", "<f4", 3), ("prev", "<u4"), ("normal", "<f4", 3), ("pdf", "<f4"),
                         ("wo", "<f4", 3), ("cos_i", "<f4"), ("wi", "<f4", 3), ("pad0", "<f4"),
                         ("throughput", "<f4", 3), ("pad1", "<f4"), ("radiance", "<f4", 3), ("pad2", "<f4")])


def bf16(x):
    """Round float32 values to bfloat16 (nearest even), returned as float32."""
    a = np.ascontiguousarray(x, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).reshape(a.shape)


def unpack(w):
    """[(W [out][in], b [out])] of the flat parameter vector."""
    w = np.asarray(w)
    out = []
    for l in range(LAYERS):
        base = l * LAYER_PARAMS
        out.append((w[base:base + DIM * DIM].reshape(DIM, DIM), w[base + DIM * DIM:base + LAYER_PARAMS]))
    return out


def features(lo, inv_extent, p, n, wo):
    """[N][64] inputs: normalised position, sin/cos(2^k pi x) for k < 6, normal, wo."""
    q = (np.asarray(p, np.float64) - np.asarray(lo, np.float64)) * np.asarray(inv_extent, np.float64)
    cols = [q]
    for a in range(3):
        for o in range(OCTAVES):
            s = np.pi * 2.0 ** o * q[:, a]
            cols.append(np.stack([np.sin(s), np.cos(s)], 1))
    cols += [np.asarray(n, np.float64), np.asarray(wo, np.float64)]
    f = np.concatenate(cols, 1)
    out = np.zeros((len(f), DIM))
    out[:, :FEATURES] = f
    return out


def forward(w, x, round=True):
    """Raw outputs [N][64] and the layer inputs [x, h1, h2, h3]."""
    layers = unpack(w)
    a = bf16(x).astype(np.float64) if round else np.asarray(x, np.float64)
    acts = [a]
    for l, (W, b) in enumerate(layers):
        Wl = bf16(W).astype(np.float64) if round else np.asarray(W, np.float64)
        z = a @ Wl.T + np.asarray(b, np.float64)
        if l == LAYERS - 1:
            return z, acts
        a = np.maximum(z, 0.0)
        if round:
            a = bf16(a.astype(np.float32)).astype(np.float64)
        acts.append(a)


def mixture(o):
    """(mu [N][8][3], kappa [N][8], pi [N][8]) of raw outputs [N][>=40]."""
    o = np.asarray(o, np.float64)[:, :OUT].reshape(-1, LOBES, 5)
    u = o[..., :3]
    ln = np.linalg.norm(u, axis=-1, keepdims=True)
    mu = np.where(ln > 1e-12, u / np.maximum(ln, 1e-12), np.array([0.0, 0.0, 1.0]))
    kappa = np.exp(np.clip(o[..., 3], LOG_KAPPA_MIN, LOG_KAPPA_MAX))
    logits = o[..., 4] - o[..., 4].max(-1, keepdims=True)
    pi = np.exp(logits)
    pi /= pi.sum(-1, keepdims=True)
    return mu, kappa, pi


def vmf_log_pdf(kappa, cos):
    return np.log(kappa) - np.log(2.0 * np.pi) - np.log1p(-np.exp(-2.0 * kappa)) + kappa * (cos - 1.0)


def log_q(o, wi):
    mu, kappa, pi = mixture(o)
    c = np.einsum("nlk,nk->nl", mu, np.asarray(wi, np.float64))
    lp = np.log(np.maximum(pi, 1e-300)) + vmf_log_pdf(kappa, c)
    m = lp.max(-1, keepdims=True)
    return (m + np.log(np.exp(lp - m).sum(-1, keepdims=True)))[:, 0]


def loss_grad(o, wi, weight):
    """d(-weight log q(wi)) / d(raw outputs), [N][40], and log q."""
    o = np.asarray(o, np.float64)[:, :OUT]
    wi = np.asarray(wi, np.float64)
    weight = np.asarray(weight, np.float64)
    mu, kappa, pi = mixture(o)
    c = np.einsum("nlk,nk->nl", mu, wi)
    lp = np.log(np.maximum(pi, 1e-300)) + vmf_log_pdf(kappa, c)
    lq = log_q(o, wi)
    r = np.exp(lp - lq[:, None])
    g = np.zeros((len(o), LOBES, 5))
    w = weight[:, None]
    g[..., 4] = -w * (r - pi)
    s = o.reshape(-1, LOBES, 5)[..., 3]
    e = np.exp(-2.0 * kappa)
    dk = 1.0 / kappa - 2.0 * e / (1.0 - e) + c - 1.0
    free = (s > LOG_KAPPA_MIN) & (s < LOG_KAPPA_MAX)
    g[..., 3] = np.where(free, -w * r * dk * kappa, 0.0)
    u = o.reshape(-1, LOBES, 5)[..., :3]
    ln = np.linalg.norm(u, axis=-1)
    sc = np.where(ln > 1e-12, -w * r * kappa / np.maximum(ln, 1e-12), 0.0)
    g[..., :3] = sc[..., None] * (wi[:, None, :] - mu * c[..., None])
    return g.reshape(-1, OUT), lq


def backward(w, acts, dout, round=True):
    """Flat gradient of sum over samples, given d loss / d outputs [N][64]."""
    layers = unpack(w)
    grad = np.zeros(PARAMS)
    d = bf16(np.asarray(dout, np.float32)).astype(np.float64) if round else np.asarray(dout, np.float64)
    for l in range(LAYERS - 1, -1, -1):
        W, _ = layers[l]
        a = acts[l]
        base = l * LAYER_PARAMS
        grad[base:base + DIM * DIM] = (d.T @ a).ravel()
        grad[base + DIM * DIM:base + LAYER_PARAMS] = d.sum(0)
        if l == 0:
            break
        Wl = bf16(W).astype(np.float64) if round else np.asarray(W, np.float64)
        d = (d @ Wl) * (a > 0.0)
        if round:
            d = bf16(d.astype(np.float32)).astype(np.float64)
    return grad


def record_weight(recs, w_max=1e30):
    li = recs["radiance"].astype(np.float64)
    lum = 0.2126 * li[:, 0] + 0.7152 * li[:, 1] + 0.0722 * li[:, 2]
    pdf = recs["pdf"].astype(np.float64)
    w = np.where(pdf > 0.0, lum * recs["cos_i"] / np.where(pdf > 0.0, pdf, 1.0), 0.0)
    return np.clip(w, 0.0, w_max)


def record_gradient(w, recs, lo, inv_extent, w_max=1e30, round=True):
    """(flat gradient, loss) of sum_i -w_i log q(wi_i) over the records."""
    x = features(lo, inv_extent, recs["position"], recs["normal"], recs["wo"])
    if round:
        x = x.astype(np.float32)
    o, acts = forward(w, x, round)
    wt = record_weight(recs, w_max)
    g, lq = loss_grad(o, recs["wi"], wt)
    dout = np.zeros((len(o), DIM))
    dout[:, :OUT] = g
    return backward(w, acts, dout, round), float(-(wt * lq).sum())


def vmf_sample(mu, kappa, u1, u2):
    """Jakob 2012: cos = 1 + log(u + (1-u) e^{-2k}) / k, azimuth 2 pi u2."""
    mu = np.asarray(mu, np.float64)
    u = np.maximum(u1, 1e-7)
    c = np.clip(1.0 + np.log(u + (1.0 - u) * np.exp(-2.0 * kappa)) / kappa, -1.0, 1.0)
    s = np.sqrt(np.maximum(0.0, 1.0 - c * c))
    phi = 2.0 * np.pi * u2
    sign = np.where(mu[:, 2] >= 0, 1.0, -1.0)
    a = -1.0 / (sign + mu[:, 2])
    b = mu[:, 0] * mu[:, 1] * a
    t = np.stack([1.0 + sign * mu[:, 0] ** 2 * a, sign * b, -sign * mu[:, 0]], 1)
    bb = np.stack([b, sign + mu[:, 1] ** 2 * a, -mu[:, 1]], 1)
    d = mu * c[:, None] + t * (s * np.cos(phi))[:, None] + bb * (s * np.sin(phi))[:, None]
    return d / np.linalg.norm(d, axis=1, keepdims=True)


def adam_step(w, g, m, v, step, lr, beta1=0.9, beta2=0.999, eps=1e-8):
    m = beta1 * m + (1.0 - beta1) * g
    v = beta2 * v + (1.0 - beta2) * g * g
    mh = m / (1.0 - beta1 ** step)
    vh = v / (1.0 - beta2 ** step)
    return w - lr * mh / (np.sqrt(vh) + eps), m, v


def init_weights(seed=0, scale=1.0):
    """Random He-uniform parameters for tests (not the library's initialiser)."""
    rng = np.random.default_rng(seed)
    w = np.zeros(PARAMS)
    for l in range(LAYERS):
        fan = FEATURES if l == 0 else DIM
        bound = np.sqrt(6.0 / fan) * scale * (0.3 if l == LAYERS - 1 else 1.0)
        base = l * LAYER_PARAMS
        W = rng.uniform(-bound, bound, (DIM, DIM))
        if l == 0:
            W[:, FEATURES:] = 0.0
        w[base:base + DIM * DIM] = W.ravel()
        w[base + DIM * DIM:base + LAYER_PARAMS] = rng.uniform(-0.1, 0.1, DIM)
    return w.astype(np.float32)
