// Learned path guiding: a small MLP maps a shading point (position,
// normal, outgoing direction) to a mixture of 8 von Mises-Fisher lobes over
// incident directions.  The renderer samples it with one-sample MIS against the
// BSDF and trains it online from the radiance its own paths carry.
//
// The reference has no path guiding; its BSDF sampling routines
// (MaterialEval.cu:270-309 diffuse, :720-836 GGX) are what the guide is
// combined with.  Published methods followed here:
//  * mixture-of-vMF output of an MLP and its KL-divergence training loss:
//    Dong et al. 2023, "Neural Parametric Mixtures for Path Guiding";
//    the loss gradient -(F/p) grad log q of Mueller et al. 2019, "Neural
//    Importance Sampling";
//  * vMF sampling in the numerically stable form of W. Jakob 2012,
//    "Numerically stable sampling of the von Mises Fisher distribution on S^2";
//  * Adam: Kingma and Ba 2015.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pt_device.hpp"

namespace pt {

constexpr int kGuideDim = 64;          // width of every layer (inputs padded)
constexpr int kGuideLayers = 4;        // 3 hidden layers + output layer
constexpr int kGuideLobes = 8;
constexpr int kGuideOut = 5 * kGuideLobes;   // per lobe: mu (3, unnormalised), log kappa, logit
constexpr int kGuideOctaves = 6;       // sin/cos(2^k pi x), k < 6, per position axis
constexpr int kGuideFeatures = 3 + 3 * 2 * kGuideOctaves + 3 + 3;   // 45
constexpr int kGuideLayerParams = kGuideDim * kGuideDim + kGuideDim; // W [out][in] then bias
constexpr int kGuideParams = kGuideLayers * kGuideLayerParams;       // 16,640
constexpr float kGuideLogKappaMin = -2.0f, kGuideLogKappaMax = 7.0f;

// Mixture parameters of one shading point: 8 lobes of (mu xyz unit, kappa,
// pi) and c = log(pi kappa / (2 pi (1 - e^{-2 kappa}))), so that the density
// is sum_j exp(c_j + kappa_j (mu_j . w - 1)).
struct GuideMixture {
    float mu[kGuideLobes][3];
    float kappa[kGuideLobes];
    float pi[kGuideLobes];
    float c[kGuideLobes];
};

// Training record written at a scatter vertex: the vertex, the sampled
// direction, the path throughput after the vertex, and the incident radiance
// gathered later along that direction (rgb, summed by the record chain walk).
struct alignas(16) GuideRecord {
    float p[3];
    uint32_t prev;          // previous record of the same path, or kInvalid
    float n[3];
    float pdf;              // density the direction was sampled with
    float wo[3];
    float cos_i;            // |n . wi|
    float wi[3];
    float pad;
    float t_after[3];       // path throughput after this vertex
    float pad2;
    float li[3];            // incident radiance estimate along wi
    float pad3;
};
static_assert(sizeof(GuideRecord) == 96, "guide record is 96 B");

// Scene box used to normalise positions, and the MIS settings.
struct GuideConfig {
    float lo[3];
    float guide_fraction;   // probability of sampling the guide (alpha)
    float inv_extent[3];
    uint32_t rec_mask;      // a path writes training records when (sample index & rec_mask) == 0
};

// ---------------------------------------------------------------- features

// Input features of a shading point, handed to put(index, value): the
// normalised position, its sin/cos frequency encoding, the normal and the
// outgoing direction (45 values; the rest of the 64 inputs are zero).
template <class Put>
__device__ void guide_features_each(const GuideConfig& g, V3 p, V3 n, V3 wo, Put put) {
    const float q[3] = {(p.x - g.lo[0]) * g.inv_extent[0], (p.y - g.lo[1]) * g.inv_extent[1],
                        (p.z - g.lo[2]) * g.inv_extent[2]};
    int k = 0;
    for (int a = 0; a < 3; a++) put(k++, q[a]);
    for (int a = 0; a < 3; a++) {
        // octave 0 directly, the others by doubling: sin 2x = 2 sin x cos x,
        // cos 2x = 1 - 2 sin^2 x
        float sn = __sinf(kPi * q[a]), cs = __cosf(kPi * q[a]);
        for (int o = 0; o < kGuideOctaves; o++) {
            put(k++, sn);
            put(k++, cs);
            const float s2 = 2.0f * sn * cs, c2 = 1.0f - 2.0f * sn * sn;
            sn = s2;
            cs = c2;
        }
    }
    put(k++, n.x);
    put(k++, n.y);
    put(k++, n.z);
    put(k++, wo.x);
    put(k++, wo.y);
    put(k++, wo.z);
}

__device__ void guide_features(const GuideConfig& g, V3 p, V3 n, V3 wo, float f[kGuideFeatures]) {
    guide_features_each(g, p, n, wo, [&](int k, float v) { f[k] = v; });
}

// ---------------------------------------------------------------- vMF mixture

// log of the vMF density kappa/(2 pi (1 - e^{-2 kappa})) e^{kappa (mu.w - 1)}.
PT_HD float vmf_log_pdf(float kappa, float mu_dot_w) {
    return logf(kappa) - logf(2.0f * kPi) - log1pf(-expf(-2.0f * kappa)) + kappa * (mu_dot_w - 1.0f);
}

// Softmax shift and normaliser of the 8 lobe logits o[5j+4].
PT_HD void guide_softmax(const float* o, float& mx, float& inv_sum) {
    mx = -1e30f;
    for (int j = 0; j < kGuideLobes; j++) mx = fmaxf(mx, o[5 * j + 4]);
    float sum = 0.0f;
    for (int j = 0; j < kGuideLobes; j++) sum += expf(o[5 * j + 4] - mx);
    inv_sum = 1.0f / sum;
}

// Lobe j of the mixture from its 5 raw outputs o5 = o + 5j.
PT_HD void guide_lobe(const float* o5, float mx, float inv_sum, float mu[3], float& kappa, float& pi, float& c) {
    const float ux = o5[0], uy = o5[1], uz = o5[2];
    const float len = sqrtf(ux * ux + uy * uy + uz * uz);
    const float inv = 1.0f / fmaxf(len, 1e-12f);
    mu[0] = len > 1e-12f ? ux * inv : 0.0f;
    mu[1] = len > 1e-12f ? uy * inv : 0.0f;
    mu[2] = len > 1e-12f ? uz * inv : 1.0f;
    kappa = expf(clampf(o5[3], kGuideLogKappaMin, kGuideLogKappaMax));
    pi = expf(o5[4] - mx) * inv_sum;
    c = logf(fmaxf(pi, 1e-30f)) + vmf_log_pdf(kappa, 1.0f);
}

// Mixture parameters from the 40 raw network outputs.
PT_HD void guide_mixture(const float o[kGuideOut], GuideMixture& m) {
    float mx, inv_sum;
    guide_softmax(o, mx, inv_sum);
    for (int j = 0; j < kGuideLobes; j++) guide_lobe(o + 5 * j, mx, inv_sum, m.mu[j], m.kappa[j], m.pi[j], m.c[j]);
}

// Mixtures of a whole bounce, packed: 8 lobes x (mu xyz, kappa, pi) as fp16,
// lobes t and t + 4 in one 32-bit word.  Word 4 c + t of entry e lies at
// w[(4 c + t) * stride + e] (c: 0-2 mu, 3 kappa, 4 pi), so that a wave reads
// and writes each word of 64 consecutive entries in one line.
constexpr int kGuideMixWords = 20;

PT_HD uint32_t guide_pack_half2(float lo, float hi) {
    const _Float16 a = (_Float16)lo, b = (_Float16)hi;
    uint16_t ua, ub;
    __builtin_memcpy(&ua, &a, 2);
    __builtin_memcpy(&ub, &b, 2);
    return (uint32_t)ua | ((uint32_t)ub << 16);
}
PT_HD float guide_half_lo(uint32_t w) {
    const uint16_t u = (uint16_t)(w & 0xFFFFu);
    _Float16 h;
    __builtin_memcpy(&h, &u, 2);
    return (float)h;
}
PT_HD float guide_half_hi(uint32_t w) {
    const uint16_t u = (uint16_t)(w >> 16);
    _Float16 h;
    __builtin_memcpy(&h, &u, 2);
    return (float)h;
}

// Lobe j of a packed mixture: unit mu (the stored fp16 vector normalised, as
// the sampler and the density both use it), kappa and the unnormalised weight.
struct GuideLobe {
    V3 mu;
    float kappa, pi;
};

// One shading point's packed mixture, read where it is used (the words stay in
// cache) so that no lobe table occupies registers across the shading code.
// The weights are renormalised to sum to one after the fp16 round trip, and the
// density is q(w) = sum_j pi_j kappa_j / (2 pi (1 - e^{-2 kappa_j})) e^{kappa_j (mu_j . w - 1)}.
// Sampling and density share make() and the hardware exp/rcp/rsqrt: both see
// the same lobes, and the density is exact to a few ulp.
struct GuideMixRef {
    const uint32_t* w;       // word 0 of this entry
    uint32_t stride;
    __device__ uint32_t word(int k) const { return w[(size_t)k * stride]; }
    __device__ static float half_of(uint32_t v, int h) { return h ? guide_half_hi(v) : guide_half_lo(v); }
    __device__ static GuideLobe make(float x, float y, float z, float k, float p) {
        GuideLobe l;
        const float len2 = x * x + y * y + z * z;
        const bool ok = len2 > 1e-12f;
        const float inv = ok ? rsqrtf(len2) : 0.0f;
        l.mu = ok ? v3(x * inv, y * inv, z * inv) : v3(0.0f, 0.0f, 1.0f);
        l.kappa = fmaxf(k, 1e-3f);
        l.pi = fmaxf(p, 0.0f);
        return l;
    }
    __device__ GuideLobe lobe(int j) const {
        const int t = j & 3, h = j >> 2;
        return make(half_of(word(t), h), half_of(word(4 + t), h), half_of(word(8 + t), h), half_of(word(12 + t), h),
                    half_of(word(16 + t), h));
    }
    __device__ float weight(int j) const { return fmaxf(half_of(word(16 + (j & 3)), j >> 2), 0.0f); }
};

__device__ __forceinline__ float guide_vmf_norm(float kappa) {
    return __fdividef(kappa, 2.0f * kPi * (1.0f - __expf(-2.0f * kappa)));
}

// The mixture folded onto the hemisphere of n: directions below the surface
// are mirrored across its plane, so the density at w (w . n >= 0) is
// q(w) + q(w - 2 (w . n) n).
PT_HD V3 guide_mirror(V3 w, V3 n) { return w - n * (2.0f * dot(w, n)); }
__device__ float guide_pdf_folded(const GuideMixRef& m, V3 w, V3 n) {
    if (dot(w, n) < 0.0f) return 0.0f;
    const V3 wr = guide_mirror(w, n);
    float q = 0.0f, sum = 0.0f;
    for (int t = 0; t < 4; t++) {
        const uint32_t x = m.word(t), y = m.word(4 + t), z = m.word(8 + t), k = m.word(12 + t), p = m.word(16 + t);
        for (int h = 0; h < 2; h++) {
            const GuideLobe l = GuideMixRef::make(GuideMixRef::half_of(x, h), GuideMixRef::half_of(y, h),
                                                  GuideMixRef::half_of(z, h), GuideMixRef::half_of(k, h),
                                                  GuideMixRef::half_of(p, h));
            const float a = l.pi * guide_vmf_norm(l.kappa);
            q += a * (__expf(l.kappa * (dot(l.mu, w) - 1.0f)) + __expf(l.kappa * (dot(l.mu, wr) - 1.0f)));
            sum += l.pi;
        }
    }
    return sum > 0.0f ? __fdividef(q, sum) : 0.0f;
}

// Orthonormal basis around a unit vector (Duff et al. 2017, "Building an
// Orthonormal Basis, Revisited").
PT_HD void guide_basis(V3 n, V3& t, V3& b) {
    const float s = copysignf(1.0f, n.z);
    const float a = -1.0f / (s + n.z);
    const float c = n.x * n.y * a;
    t = v3(1.0f + s * n.x * n.x * a, s * c, -s * n.x);
    b = v3(c, s + n.y * n.y * a, -n.y);
}

// Direction from the mixture: u0 picks the lobe (weights renormalised as in
// the density), (u1, u2) the direction within it (Jakob 2012:
// cos = 1 + log(u + (1-u) e^{-2 kappa}) / kappa).
__device__ V3 guide_sample(const GuideMixRef& m, float u0, float u1, float u2) {
    float sum = 0.0f;
    for (int j = 0; j < kGuideLobes; j++) sum += m.weight(j);
    const float target = u0 * sum;
    int j = 0;
    float acc = m.weight(0);
    while (j < kGuideLobes - 1 && target >= acc) {
        j++;
        acc += m.weight(j);
    }
    const GuideLobe l = m.lobe(j);
    const float k = l.kappa;
    const float u = fmaxf(u1, 1e-7f);
    const float c = clampf(1.0f + __fdividef(__logf(u + (1.0f - u) * __expf(-2.0f * k)), k), -1.0f, 1.0f);
    const float s = sqrtf(fmaxf(0.0f, 1.0f - c * c));
    const float phi = 2.0f * kPi * u2;
    V3 t, b;
    guide_basis(l.mu, t, b);
    return normalize(l.mu * c + t * (s * __cosf(phi)) + b * (s * __sinf(phi)));
}

__device__ V3 guide_sample_folded(const GuideMixRef& m, V3 n, float u0, float u1, float u2) {
    const V3 w = guide_sample(m, u0, u1, u2);
    return dot(w, n) < 0.0f ? guide_mirror(w, n) : w;
}

// d(loss)/d(raw outputs) of one record for loss = -weight * log q(wi).
// Returns log q(wi).
PT_HD float guide_loss_grad(const float o[kGuideOut], V3 wi, float weight, float g[kGuideOut]) {
    GuideMixture m;
    guide_mixture(o, m);
    float lp[kGuideLobes];
    float mxl = -1e30f;
    for (int j = 0; j < kGuideLobes; j++) {
        const float d = m.mu[j][0] * wi.x + m.mu[j][1] * wi.y + m.mu[j][2] * wi.z;
        lp[j] = logf(fmaxf(m.pi[j], 1e-30f)) + vmf_log_pdf(m.kappa[j], d);
        mxl = fmaxf(mxl, lp[j]);
    }
    float sum = 0.0f;
    for (int j = 0; j < kGuideLobes; j++) sum += expf(lp[j] - mxl);
    const float logq = mxl + logf(sum);
    for (int j = 0; j < kGuideLobes; j++) {
        const float r = expf(lp[j] - logq);            // responsibility of lobe j
        // logits: d log q / d logit_j = r_j - pi_j
        g[5 * j + 4] = -weight * (r - m.pi[j]);
        // kappa = exp(s) (clamped): d log vMF / d kappa = 1/kappa - 2e^{-2k}/(1-e^{-2k}) + mu.w - 1
        const float s = o[5 * j + 3];
        const float k = m.kappa[j];
        const float d = m.mu[j][0] * wi.x + m.mu[j][1] * wi.y + m.mu[j][2] * wi.z;
        const float e = expf(-2.0f * k);
        const float dk = 1.0f / k - 2.0f * e / (1.0f - e) + d - 1.0f;
        const bool free_k = s > kGuideLogKappaMin && s < kGuideLogKappaMax;
        g[5 * j + 3] = free_k ? -weight * r * dk * k : 0.0f;
        // mu = u/|u|: d log vMF / d u = kappa (I - mu mu^T) wi / |u|
        const float ux = o[5 * j], uy = o[5 * j + 1], uz = o[5 * j + 2];
        const float len = sqrtf(ux * ux + uy * uy + uz * uz);
        const float sc = len > 1e-12f ? -weight * r * k / len : 0.0f;
        g[5 * j + 0] = sc * (wi.x - m.mu[j][0] * d);
        g[5 * j + 1] = sc * (wi.y - m.mu[j][1] * d);
        g[5 * j + 2] = sc * (wi.z - m.mu[j][2] * d);
    }
    return logq;
}

}  // namespace pt
