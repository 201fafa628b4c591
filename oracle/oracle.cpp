// ORACLE - test infrastructure only.
//
// A CPU restatement of the reference's wavefront path-tracing stages, used by
// tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py as the
// checker of the GPU core.  Nothing in the product links, loads or calls it.
//
// What it restates (reference files under KrzystekMatej/gpu-path-tracer):
//   * HostBvh object-median BVH2 + DFS flatten      Core/Source/Graphics/Cuda/Bvh/HostBvh.cpp:22-100,
//                                                    Core/Source/Graphics/Cuda/Bvh/DeviceBvh.cpp:7-52
//   * slab test, Moller-Trumbore, closest/any hit   Core/Source/Graphics/Cuda/PathTracing/Kernels/Intersect.cu:11-179
//   * alias table build / sample                    .../PathTracing/Memory/AliasTable.hpp:22-84,
//                                                    .../PathTracing/Memory/AliasTableView.hpp:32-60
//   * light sampler                                 Core/Source/Graphics/Cuda/PathTracing/Memory/LightSampler.cpp:5-40
//   * camera and path generation                    Core/Source/Graphics/Cuda/PathTracing/Renderer.cpp:237-257,
//                                                    .../Kernels/PathGeneration.cu:15-102
//   * diffuse / GGX / mirror / emissive, NEE, RR    .../Kernels/MaterialEval.cu:25-392,587-912
//   * shadow clamp and accumulation                 .../Kernels/Intersect.cu:125-179, .../Kernels/Common.hpp:8-16
//   * tonemap                                       .../Kernels/Image.cu:34-54
// plus PCG32 (O'Neill 2014) and Brown's LCG jump-ahead for the sample streams,
// and a brute-force closest hit under the (t, prim id) order.
//
// Pinning: tests/test_oracle.py checks each function against a closed form,
// brute force or a published known-answer vector.  Parity unpinned: the
// cuRAND Philox stream and CUDA texture filtering of the reference (not
// reproducible here), see the oracle section of the design notes.
//
// Built with g++ -O2 -ffp-contract=off: float expressions round exactly as
// written, as they do in the GPU build.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <memory>
#include <thread>
#include <vector>

namespace orc {

constexpr uint32_t kInvalid = 0xFFFFFFFFu;
constexpr float kMinT = 1e-4f;
constexpr float kMaxT = 1e34f;
constexpr float kPi = 3.14159265358979323846f;
constexpr float kInvPi = 1.0f / 3.14159265358979323846f;
constexpr float kOneMinusEpsilon = 0x1.fffffep-1f;

enum Model : uint32_t { Normal = 0, Diffuse = 1, Mirror = 2, Phong = 3, Ggx = 4, Emissive = 5 };

// ------------------------------------------------------------------ float3 (helper_math.h host semantics)
struct F3 {
    float x, y, z;
};
inline F3 f3(float x, float y, float z) { return F3{x, y, z}; }
inline F3 f3(float s) { return F3{s, s, s}; }
inline F3 operator+(F3 a, F3 b) { return f3(a.x + b.x, a.y + b.y, a.z + b.z); }
inline F3 operator-(F3 a, F3 b) { return f3(a.x - b.x, a.y - b.y, a.z - b.z); }
inline F3 operator-(F3 a) { return f3(-a.x, -a.y, -a.z); }
inline F3 operator*(F3 a, F3 b) { return f3(a.x * b.x, a.y * b.y, a.z * b.z); }
inline F3 operator*(F3 a, float s) { return f3(a.x * s, a.y * s, a.z * s); }
inline F3 operator*(float s, F3 a) { return f3(s * a.x, s * a.y, s * a.z); }
inline F3 operator/(F3 a, float s) { return f3(a.x / s, a.y / s, a.z / s); }
inline F3 operator/(F3 a, F3 b) { return f3(a.x / b.x, a.y / b.y, a.z / b.z); }
inline float dot(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline F3 cross(F3 a, F3 b) { return f3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
inline float length(F3 a) { return std::sqrt(dot(a, a)); }
// helper_math.h:78-81 host rsqrtf = 1/sqrtf; normalize = v * rsqrt(v.v) (:1325-1329)
inline float rsqrt_host(float x) { return 1.0f / std::sqrt(x); }
inline F3 normalize(F3 v) { return v * rsqrt_host(dot(v, v)); }
inline F3 reflect(F3 i, F3 n) { return i - 2.0f * n * dot(n, i); }   // helper_math.h:1427-1430
inline F3 fmin3(F3 a, F3 b) { return f3(std::fmin(a.x, b.x), std::fmin(a.y, b.y), std::fmin(a.z, b.z)); }
inline F3 fmax3(F3 a, F3 b) { return f3(std::fmax(a.x, b.x), std::fmax(a.y, b.y), std::fmax(a.z, b.z)); }
inline float max_component(F3 v) { return std::fmax(v.x, std::fmax(v.y, v.z)); }   // Math.hpp:81-84
inline float min_component(F3 v) { return std::fmin(v.x, std::fmin(v.y, v.z)); }
inline float clampf(float v, float a, float b) { return std::fmax(a, std::fmin(v, b)); }   // helper_math.h:1168
inline F3 lerp(F3 a, F3 b, float t) { return a + t * (b - a); }   // helper_math.h:1146-1161
inline float sqr(float v) { return v * v; }
inline float at(F3 v, int i) { return i == 0 ? v.x : i == 1 ? v.y : v.z; }

// ------------------------------------------------------------------ PCG32 (pcg-c-basic restated)
constexpr uint64_t kMult = 6364136223846793005ull;
struct Pcg {
    uint64_t state, inc;
};
inline uint32_t pcg_next(Pcg& r) {
    const uint64_t old = r.state;
    r.state = old * kMult + r.inc;
    const uint32_t xorshifted = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    const uint32_t rot = (uint32_t)(old >> 59u);
    return (xorshifted >> rot) | (xorshifted << ((-rot) & 31));
}
inline Pcg pcg_srandom(uint64_t initstate, uint64_t initseq) {
    Pcg r{0u, (initseq << 1u) | 1u};
    pcg_next(r);
    r.state += initstate;
    pcg_next(r);
    return r;
}
// Brown, "Random Number Generation with Arbitrary Stride" (1994)
inline void pcg_advance(Pcg& r, uint64_t delta) {
    uint64_t cur_mult = kMult, cur_plus = r.inc, acc_mult = 1u, acc_plus = 0u;
    while (delta > 0) {
        if (delta & 1) {
            acc_mult *= cur_mult;
            acc_plus = acc_plus * cur_mult + cur_plus;
        }
        cur_plus = (cur_mult + 1) * cur_plus;
        cur_mult *= cur_mult;
        delta /= 2;
    }
    r.state = acc_mult * r.state + acc_plus;
}
inline float to_unit(uint32_t x) { return (float)(x >> 8) * 0x1p-24f; }

// Sample stream: draws 0,1 camera jitter; bounce b owns draws 2+8b .. 9+8b.
struct SampleRng {
    Pcg r;
    SampleRng(uint64_t seed, uint64_t sample, uint64_t first_draw) {
        r = pcg_srandom(seed, sample);
        pcg_advance(r, first_draw);
    }
    float next() { return to_unit(pcg_next(r)); }
};

// ------------------------------------------------------------------ scene records
struct Material {          // pt_material of the C ABI
    uint32_t model;
    float color[3];
    float specular[3];
    float roughness;
    float metallic;
    float ior;
    float emission[3];
    float normal_ts[3];
};

struct Triangle {          // Triangle.hpp:8-38 (intersection + shading + positions)
    F3 e1, e2, v0;
    uint32_t model;
    F3 n[3];
    float t[3][4];
    F3 gn;
    uint32_t material;
    F3 p[3];
    uint32_t prim;         // index in the caller's (pre-BVH) order
};

struct Light {             // LightTriangle, Triangle.hpp:40-58
    F3 e1, e2, v0;
    uint32_t index;
    F3 gn;
    float area;
    F3 emission;
};

struct Bin {               // Bin<uint32_t>, AliasTableView.hpp:8-14
    float q, p;
    uint32_t alias;
    uint32_t value;
};

struct Box {
    F3 min, max;
};

struct Node {              // DeviceBvhNode, DeviceBvhNode.hpp:8-23
    Box bounds;
    uint32_t left, right, first, count;
    bool leaf() const { return left == kInvalid && right == kInvalid; }
};

// ------------------------------------------------------------------ alias table
std::vector<Bin> build_alias(const std::vector<float>& weights, const std::vector<uint32_t>& values) {
    double s = 0.0;   // std::accumulate(..., 0.0)
    for (float w : weights) s = s + w;
    const float sum = (float)s;
    std::vector<Bin> bins(weights.size());
    for (uint32_t i = 0; i < weights.size(); i++) bins[i] = Bin{0.0f, weights[i] / sum, 0u, values[i]};
    struct Outcome {
        float pHat;
        uint32_t index;
    };
    std::vector<Outcome> under, over;
    for (uint32_t i = 0; i < bins.size(); i++) {
        const float pHat = bins[i].p * bins.size();
        if (pHat < 1) under.push_back(Outcome{pHat, i});
        else over.push_back(Outcome{pHat, i});
    }
    while (!under.empty() && !over.empty()) {
        Outcome un = under.back(), ov = over.back();
        under.pop_back();
        over.pop_back();
        bins[un.index].q = un.pHat;
        bins[un.index].alias = ov.index;
        const float pExcess = un.pHat + ov.pHat - 1;
        if (pExcess < 1) under.push_back(Outcome{pExcess, ov.index});
        else over.push_back(Outcome{pExcess, ov.index});
    }
    while (!over.empty()) {
        bins[over.back().index].q = 1;
        bins[over.back().index].alias = kInvalid;
        over.pop_back();
    }
    while (!under.empty()) {
        bins[under.back().index].q = 1;
        bins[under.back().index].alias = kInvalid;
        under.pop_back();
    }
    return bins;
}

inline uint32_t alias_sample(const Bin* bins, uint32_t size, float u, float* pmf) {
    const float scaledU = u * static_cast<float>(size);
    const uint32_t binIndex = std::min(static_cast<uint32_t>(scaledU), size - 1u);
    const float up = std::fmin(scaledU - static_cast<float>(binIndex), kOneMinusEpsilon);
    const Bin bin = bins[binIndex];
    if (up < bin.q) {
        *pmf = bin.p;
        return binIndex;
    }
    *pmf = bins[bin.alias].p;
    return bin.alias;
}

// ------------------------------------------------------------------ scene + HostBvh
struct Scene {
    std::vector<Triangle> tris;      // BVH (post nth_element) order
    std::vector<Material> materials;
    std::vector<Light> lights;
    std::vector<Bin> bins;
    std::vector<Node> nodes;
    std::vector<uint32_t> by_prim;   // prim id -> index in tris
    uint32_t depth = 0, node_count = 0, max_depth = 0, min_triangles = 0;
};

uint32_t estimate_leaf(size_t n) {   // HostBvh.cpp:22-30
    if (n < 5'000) return 4;
    if (n < 50'000) return 6;
    if (n < 500'000) return 8;
    if (n < 2'000'000) return 12;
    if (n < 10'000'000) return 16;
    return 24;
}
uint32_t estimate_depth(size_t n) {   // HostBvh.cpp:32-35
    return std::min(64u, static_cast<uint32_t>(std::max(1.0, std::ceil(2 * std::log2((double)n)))));
}
Box pad_bounds(const Box& b) {   // HostBvh.cpp:37-48
    const F3 extent = b.max - b.min;
    const float largest = std::fmax(extent.x, std::fmax(extent.y, extent.z));
    const float padding = std::fmax(1e-5f * largest, 1e-5f);
    return Box{b.min - f3(padding), b.max + f3(padding)};
}

struct HostNode {
    Box bounds;
    uint32_t first = 0, count = 0;
    std::unique_ptr<HostNode> left, right;
};

std::unique_ptr<HostNode> build_recurse(Scene& s, uint32_t first, uint32_t count, uint32_t depth) {
    s.node_count++;
    s.depth = std::max(s.depth, depth);
    auto node = std::make_unique<HostNode>();
    node->first = first;
    node->count = count;
    const float fmax = std::numeric_limits<float>::max();
    Box b{f3(fmax), f3(-fmax)};
    for (uint32_t i = first; i < first + count; i++)
        for (const F3& p : s.tris[i].p) {
            b.min = f3(std::min(b.min.x, p.x), std::min(b.min.y, p.y), std::min(b.min.z, p.z));
            b.max = f3(std::max(b.max.x, p.x), std::max(b.max.y, p.y), std::max(b.max.z, p.z));
        }
    node->bounds = count == 0 ? Box{f3(fmax), f3(-fmax)} : pad_bounds(b);
    if (depth >= s.max_depth || count <= s.min_triangles) return node;
    uint32_t axis = 0;
    const float ex = b.max.x - b.min.x, ey = b.max.y - b.min.y, ez = b.max.z - b.min.z;
    if (ey > ex && ey > ez) axis = 1;
    else if (ez > ex && ez > ey) axis = 2;
    const uint32_t mid = first + count / 2;
    std::nth_element(s.tris.begin() + first, s.tris.begin() + mid, s.tris.begin() + first + count,
                     [axis](const Triangle& a, const Triangle& c) {
                         const F3 ca = (a.p[0] + a.p[1] + a.p[2]) / 3.0f;
                         const F3 cc = (c.p[0] + c.p[1] + c.p[2]) / 3.0f;
                         return at(ca, axis) < at(cc, axis);
                     });
    node->left = build_recurse(s, first, mid - first, depth + 1);
    node->right = build_recurse(s, mid, count - (mid - first), depth + 1);
    return node;
}

uint32_t flatten(const HostNode& n, std::vector<Node>& out) {
    const uint32_t idx = (uint32_t)out.size();
    out.push_back(Node{n.bounds, kInvalid, kInvalid, n.first, n.count});
    if (n.left && n.right) {
        const uint32_t l = flatten(*n.left, out);
        out[idx].left = l;
        const uint32_t r = flatten(*n.right, out);
        out[idx].right = r;
    }
    return idx;
}

// ------------------------------------------------------------------ intersection (Intersect.cu)
struct Ray {
    F3 o, d;
    float tmin, tmax;
};

inline bool aabb_intersect(const Box& b, const Ray& r, F3 inv) {   // Intersect.cu:11-23
    const F3 tn = (b.min - r.o) * inv;
    const F3 tf = (b.max - r.o) * inv;
    const F3 lo = fmin3(tn, tf), hi = fmax3(tn, tf);
    const float t0 = std::fmax(r.tmin, max_component(lo));
    const float t1 = std::fmin(r.tmax, min_component(hi));
    return t0 <= t1;
}

// Intersect.cu:25-61 without the upper bound (callers apply their own)
inline bool mt(const Triangle& tr, F3 o, F3 d, float tmin, float& t, float& u, float& v) {
    const F3 p = cross(d, tr.e2);
    const float det = dot(tr.e1, p);
    if (std::fabs(det) < 1e-7f) return false;
    const float invDet = 1.0f / det;
    const F3 s = o - tr.v0;
    u = dot(s, p) * invDet;
    if (u < 0.0f || u > 1.0f) return false;
    const F3 q = cross(s, tr.e1);
    v = dot(d, q) * invDet;
    if (v < 0.0f || u + v > 1.0f) return false;
    t = dot(tr.e2, q) * invDet;
    return t >= tmin;
}

struct HitRec {
    uint32_t tri;    // BVH-order index, kInvalid = miss
    float t, u, v;
};

// IntersectRaysWithSceneKernel, Intersect.cu:63-123: first-found, strict t < tMax,
// left pushed before right (right popped first).
HitRec closest_ref(const Scene& s, Ray ray, uint32_t* nodes_visited, uint32_t* tris_tested) {
    HitRec h{kInvalid, 0.0f, 0.0f, 0.0f};
    const F3 inv = f3(1.0f) / ray.d;
    uint32_t stack[70];
    uint32_t sp = 0;
    stack[sp++] = 0;
    uint32_t nv = 0, nt = 0;
    while (sp) {
        const Node& node = s.nodes[stack[--sp]];
        nv++;
        if (!aabb_intersect(node.bounds, ray, inv)) continue;
        if (node.leaf()) {
            for (uint32_t i = 0; i < node.count; i++) {
                float t, u, v;
                nt++;
                if (mt(s.tris[node.first + i], ray.o, ray.d, ray.tmin, t, u, v) && t < ray.tmax) {
                    ray.tmax = t;
                    h = HitRec{node.first + i, t, u, v};
                }
            }
        } else {
            stack[sp++] = node.left;
            stack[sp++] = node.right;
        }
    }
    if (nodes_visited) *nodes_visited = nv;
    if (tris_tested) *tris_tested = nt;
    return h;
}

// IntersectShadowRaysWithSceneKernel occlusion test, Intersect.cu:136-174
bool any_ref(const Scene& s, const Ray& ray) {
    const F3 inv = f3(1.0f) / ray.d;
    uint32_t stack[70];
    uint32_t sp = 0;
    stack[sp++] = 0;
    while (sp) {
        const Node& node = s.nodes[stack[--sp]];
        if (!aabb_intersect(node.bounds, ray, inv)) continue;
        if (node.leaf()) {
            for (uint32_t i = 0; i < node.count; i++) {
                float t, u, v;
                if (mt(s.tris[node.first + i], ray.o, ray.d, ray.tmin, t, u, v) && t < ray.tmax) return true;
            }
        } else {
            stack[sp++] = node.left;
            stack[sp++] = node.right;
        }
    }
    return false;
}

// Brute force over all triangles, minimum under the (t, prim id) order.
// The reference's traversal (closest_ref) with the acceptance test of
// closest_brute: a triangle at exactly the best t replaces it when its
// primitive id is lower.  The slab test keeps every box whose entry is at or
// before the best t, so this returns the (t, prim id) minimum of all
// triangles, as brute force does, in the time of a BVH walk.
HitRec closest_tprim(const Scene& s, Ray ray) {
    HitRec h{kInvalid, 0.0f, 0.0f, 0.0f};
    uint32_t best_prim = kInvalid;
    const F3 inv = f3(1.0f) / ray.d;
    uint32_t stack[70];
    uint32_t sp = 0;
    stack[sp++] = 0;
    while (sp) {
        const Node& node = s.nodes[stack[--sp]];
        if (!aabb_intersect(node.bounds, ray, inv)) continue;
        if (node.leaf()) {
            for (uint32_t i = 0; i < node.count; i++) {
                float t, u, v;
                if (!mt(s.tris[node.first + i], ray.o, ray.d, ray.tmin, t, u, v)) continue;
                const uint32_t prim = s.tris[node.first + i].prim;
                if (t < ray.tmax || (t == ray.tmax && best_prim != kInvalid && prim < best_prim)) {
                    ray.tmax = t;
                    best_prim = prim;
                    h = HitRec{node.first + i, t, u, v};
                }
            }
        } else {
            stack[sp++] = node.left;
            stack[sp++] = node.right;
        }
    }
    return h;
}

HitRec closest_brute(const Scene& s, const Ray& ray) {
    HitRec h{kInvalid, 0.0f, 0.0f, 0.0f};
    uint32_t best_prim = kInvalid;
    float best_t = ray.tmax;
    for (uint32_t i = 0; i < s.tris.size(); i++) {
        float t, u, v;
        if (!mt(s.tris[i], ray.o, ray.d, ray.tmin, t, u, v)) continue;
        const uint32_t prim = s.tris[i].prim;
        if (t < best_t || (t == best_t && best_prim != kInvalid && prim < best_prim)) {
            best_t = t;
            best_prim = prim;
            h = HitRec{i, t, u, v};
        }
    }
    return h;
}

bool any_brute(const Scene& s, const Ray& ray) {
    for (const Triangle& tr : s.tris) {
        float t, u, v;
        if (mt(tr, ray.o, ray.d, ray.tmin, t, u, v) && t < ray.tmax) return true;
    }
    return false;
}

// ------------------------------------------------------------------ shading (MaterialEval.cu)
struct Frame {
    F3 t, b, n;
    F3 to_local(F3 d) const { return f3(dot(d, t), dot(d, b), dot(d, n)); }
    F3 from_local(F3 d) const { return d.x * t + d.y * b + d.z * n; }
};

F3 safe_normalize(F3 v, F3 fallback) {   // Math.hpp:108-118
    const float l2 = dot(v, v);
    if (l2 > 1e-16f) return v * rsqrt_host(l2);
    return fallback;
}

Frame build_frame(F3 n) {   // Math.hpp:120-134
    F3 t;
    if (std::fabs(n.x) > std::fabs(n.z)) t = normalize(f3(n.y, -n.x, 0.0f));
    else t = normalize(f3(0.0f, n.z, -n.y));
    return Frame{t, cross(n, t), n};
}

Frame build_frame(F3 n, F3 t, float sign) {   // Math.hpp:136-148
    n = safe_normalize(n, f3(0.0f, 0.0f, 1.0f));
    const Frame fb = build_frame(n);
    t = t - n * dot(n, t);
    t = safe_normalize(t, fb.t);
    return Frame{t, cross(n, t) * sign, n};
}

F3 fix_normal(F3 gn, F3 d) { return dot(gn, d) > 0.0f ? -gn : gn; }   // MaterialEval.cu:25-30
float luminance(F3 c) { return 0.2126f * c.x + 0.7152f * c.y + 0.0722f * c.z; }

// sin and cos of a sampling angle x = 2 pi u, 0 <= x < 2 pi.  The reference
// calls the device libm here (MaterialEval.cu:228,591), whose last-ulp
// behaviour is not reproducible off NVIDIA hardware; the core and this oracle
// both evaluate the same published method instead, operation for operation:
// quadrant reduction by pi/2 split into three parts (Cody & Waite 1980; the
// first two parts have 17 significant bits, so q * part is exact for q <= 4,
// and x - q * P1 is exact by Sterbenz's lemma) and the Taylor polynomials of
// degree 9 (sin) and 10 (cos) on |r| <= pi/4 (truncation below 2e-9).
// tests/test_oracle.py bounds its error against float64 sin/cos.
void sincos_angle(float x, float* s, float* c) {
    constexpr float kTwoOverPi = 0x1.45f306p-1f;
    constexpr float kP1 = 0x1.921ep+0f, kP2 = 0x1.b544p-16f, kP3 = 0x1.0b4612p-34f;
    const float q = std::nearbyint(x * kTwoOverPi);
    float r = x - q * kP1;
    r = r - q * kP2;
    r = r - q * kP3;
    const float r2 = r * r;
    const float sp = r + r * (r2 * (-0x1.555556p-3f + r2 * (0x1.111112p-7f + r2 * (-0x1.a01a02p-13f + r2 * 0x1.71de3ap-19f))));
    const float cp = (1.0f - 0.5f * r2) +
                     (r2 * r2) * (0x1.555556p-5f + r2 * (-0x1.6c16c2p-10f + r2 * (0x1.a01a02p-16f + r2 * -0x1.27e4fcp-22f)));
    switch ((int)q & 3) {
        case 0: *s = sp; *c = cp; break;
        case 1: *s = cp; *c = -sp; break;
        case 2: *s = -sp; *c = -cp; break;
        default: *s = -cp; *c = sp; break;
    }
}

F3 cosine_hemisphere(float u0, float u1) {   // MaterialEval.cu:222-229
    const float phi = 2.0f * kPi * u0;
    const float cosTheta = std::sqrt(u1);
    const float sinTheta = std::sqrt(1.0f - u1);
    float sp, cp;
    sincos_angle(phi, &sp, &cp);
    return f3(sinTheta * cp, sinTheta * sp, cosTheta);
}

float cos2t(F3 w) { return w.z * w.z; }
float sin2t(F3 w) { return std::fmax(0.0f, 1.0f - cos2t(w)); }
float sint(F3 w) { return std::sqrt(sin2t(w)); }
float tan2t(F3 w) { return sin2t(w) / cos2t(w); }
float cosphi(F3 w) {
    const float st = sint(w);
    return st == 0 ? 1 : clampf(w.x / st, -1.f, 1.f);
}
float sinphi(F3 w) {
    const float st = sint(w);
    return st == 0 ? 0 : clampf(w.y / st, -1.f, 1.f);
}
float lambda(F3 w, float alpha) {   // MaterialEval.cu:612-619
    const float t2 = tan2t(w);
    if (t2 == std::numeric_limits<float>::infinity()) return 0.0f;
    const float a2 = sqr(cosphi(w) * alpha) + sqr(sinphi(w) * alpha);
    return (std::sqrt(1.0f + a2 * t2) - 1.0f) * 0.5f;
}
float g1(F3 w, float alpha) { return 1.0f / (1.0f + lambda(w, alpha)); }
float gsmith(F3 wo, F3 wi, float alpha) { return 1 / (1 + lambda(wo, alpha) + lambda(wi, alpha)); }
float dggx(F3 wm, float alpha) {   // MaterialEval.cu:631-640
    const float t2 = tan2t(wm);
    if (t2 == std::numeric_limits<float>::infinity()) return 0.0f;
    const float cos4 = sqr(cos2t(wm));
    if (cos4 < 1e-16f) return 0.0f;
    const float e = t2 * (sqr(cosphi(wm) / alpha) + sqr(sinphi(wm) / alpha));
    return 1 / (kPi * sqr(alpha) * cos4 * sqr(1 + e));
}
float dvis(F3 w, F3 wm, float alpha) { return g1(w, alpha) / std::fabs(w.z) * dggx(wm, alpha) * std::fabs(dot(w, wm)); }
// FresnelSchlick, MaterialEval.cu:647-650: (1 - c)^5 as three products in
// place of powf (same value up to rounding; identical on the core and here).
F3 fresnel(F3 f0, float c) {
    const float x = 1.0f - c;
    const float x2 = x * x;
    return f0 + (f3(1.0f) - f0) * ((x2 * x2) * x);
}
F3 sample_wm(F3 w, float alpha, float u0, float u1) {   // MaterialEval.cu:587-610
    F3 wh = normalize(f3(alpha * w.x, alpha * w.y, w.z));
    if (wh.z < 0) wh = -wh;
    const F3 t1 = (wh.z < 0.999f) ? normalize(cross(wh, f3(0.0f, 0.0f, 1.0f))) : f3(1.0f, 0.0f, 0.0f);
    const F3 t2 = cross(wh, t1);
    const float r = std::sqrt(u0);
    const float theta = 2.0f * kPi * u1;
    float st, ct;
    sincos_angle(theta, &st, &ct);
    float px = r * ct, py = r * st;
    const float h = std::sqrt(1.0f - px * px);
    py = h + ((1 + wh.z) * 0.5f) * (py - h);
    const float pz = std::sqrt(std::fmax(0.0f, 1.0f - (px * px + py * py)));
    const F3 nh = px * t1 + py * t2 + pz * wh;
    return normalize(f3(alpha * nh.x, alpha * nh.y, std::fmax(1e-6f, nh.z)));
}

// GGX constants of one hit, MaterialEval.cu:699-701,747-750,860-861:
// alpha = max(roughness^2, 1e-3), F0 = lerp(((eta_i - eta)/(eta_i + eta))^2, base, metallic),
// diffuse colour base * (1 - metallic).
struct GgxParams {
    F3 f0, diffc;
    float alpha;
};
GgxParams ggx_params(const Material& m, float cur_ior) {
    const F3 basec = f3(m.color[0], m.color[1], m.color[2]);
    const float metallic = m.metallic;
    const float df0 = sqr((cur_ior - m.ior) / (cur_ior + m.ior));
    return GgxParams{lerp(f3(df0), basec, metallic), basec * (1.0f - metallic), std::fmax(m.roughness * m.roughness, 0.001f)};
}

// BSDF of a light sample, EvaluateGgxDirect, MaterialEval.cu:652-718 (local
// directions): (1 - F) diffuse / pi + F D G / (4 cos_i cos_o); false where the
// reference rejects the sample (either direction below the surface, a half
// vector below it, or wo.wm / wi.wm <= 1e-6).
bool ggx_direct_bsdf(const GgxParams& g, F3 wo, F3 wl, F3* f) {
    const float ci = wl.z, co = wo.z;
    if (ci <= 0.0f || co <= 0.0f) return false;
    const F3 wm = normalize(wo + wl);
    if (wm.z <= 0.0f) return false;
    const float owm = dot(wo, wm), iwm = dot(wl, wm);
    if (owm <= 1e-6f || iwm <= 1e-6f) return false;
    const F3 F = fresnel(g.f0, clampf(owm, 0.0f, 1.0f));
    const F3 fd = (f3(1.0f) - F) * g.diffc * kInvPi;
    const F3 fs = F * dggx(wm, g.alpha) * gsmith(wo, wl, g.alpha) / (4.0f * ci * co);
    *f = fd + fs;
    return true;
}

// Lobe probabilities of the indirect sample, MaterialEval.cu:752-760: by the
// luminance of the diffuse colour and of F0; false if both are ~0.
bool ggx_lobes(const GgxParams& g, float* pd, float* ps) {
    const float dw = luminance(g.diffc), sw = luminance(g.f0);
    const float ws = dw + sw;
    if (ws <= 1e-6f) return false;
    *pd = dw / ws;
    *ps = sw / ws;
    return true;
}

// Solid-angle density of the indirect sample's direction wl (z > 0):
// pd cos / pi + ps D_wo(wm) / (4 wo.wm), MaterialEval.cu:805-807.
float ggx_pdf(float pd, float ps, F3 wo, F3 wl, F3 wm, float owm, float alpha) {
    return pd * (wl.z * kInvPi) + ps * (dvis(wo, wm, alpha) / (4.0f * owm));
}

// EvaluateGgxIndirect, MaterialEval.cu:720-836, in local coordinates: pick a
// lobe with u2, sample a cosine direction or reflect wo about a visible
// normal, then the BSDF, the mixture density and the weight f cos / pdf.
// false where the reference drops the path (before its world-space
// geometric-normal test, which the caller applies).
struct GgxSample {
    F3 wl, f, weight;
    float pdf;
};
bool ggx_indirect(const GgxParams& g, F3 wo, float u0, float u1, float u2, GgxSample* out) {
    if (!(wo.z > 0.0f)) return false;
    float pd, ps;
    if (!ggx_lobes(g, &pd, &ps)) return false;
    F3 wl;
    if (u2 < pd) wl = cosine_hemisphere(u0, u1);
    else wl = reflect(-wo, sample_wm(wo, g.alpha, u0, u1));
    const float ci = wl.z, co = wo.z;
    if (ci <= 0.0f) return false;
    const F3 wm = normalize(wl + wo);
    if (wm.z <= 0.0f) return false;
    const float owm = dot(wo, wm);
    if (owm <= 1e-6f) return false;
    const F3 F = fresnel(g.f0, clampf(owm, 0.0f, 1.0f));
    const F3 fd = (f3(1.0f) - F) * g.diffc * kInvPi;
    const F3 fs = F * dggx(wm, g.alpha) * gsmith(wo, wl, g.alpha) / (4.0f * ci * co);
    const float pdf = ggx_pdf(pd, ps, wo, wl, wm, owm, g.alpha);
    if (pdf <= 1e-8f) return false;
    out->wl = wl;
    out->f = fd + fs;
    out->pdf = pdf;
    out->weight = (fd + fs) * (ci / pdf);
    return true;
}

struct RenderParams {
    uint32_t width, height, spp, axis, depth, use_nee, rr_start;
    uint32_t tprim;   // 1: closest hits under the (t, prim id) order (the core's), 0: the reference's order
    uint64_t seed;
    F3 origin, llc, horizontal, vertical;
};

// One traced ray of a logged sample: kind 0 = extension ray (result: prim id
// bits, t, u, v), 1 = shadow ray (result: occluded, radiance before the clamp).
struct PathRecord {
    float kind, o[3], tmin, d[3], tmax, result[4], depth, pad[2];
};

struct RenderCounters {
    uint64_t ext = 0, shadow = 0;
    std::vector<PathRecord>* log = nullptr;
};

// (radiance of one sample) Follows one sample index through the wavefront
// stages; the order in which the reference's wavefront schedules samples does
// not change any sample's value because every draw is addressed by
// (sample, bounce, slot).
F3 trace_sample(const Scene& s, const RenderParams& rp, uint64_t sample, RenderCounters& rc, uint32_t* pixel_out) {
    // InitializePathsKernel / ComputePathSample / GenerateCameraRay, PathGeneration.cu:15-76
    SampleRng cam(rp.seed, sample, 0);
    const float jx = cam.next(), jy = cam.next();
    const uint64_t pix64 = sample / rp.spp;
    const uint32_t cell = (uint32_t)(sample - pix64 * rp.spp);
    const uint32_t pixel = (uint32_t)pix64;
    const uint32_t x = pixel % rp.width, y = pixel / rp.width;
    const uint32_t cx = cell % rp.axis, cy = cell / rp.axis;
    const float u = (x + (cx + jx) / rp.axis) / rp.width;
    const float v = (y + (cy + jy) / rp.axis) / rp.height;
    *pixel_out = pixel;
    Ray ray{rp.origin, normalize(rp.llc + u * rp.horizontal + v * rp.vertical - rp.origin), kMinT, kMaxT};
    F3 thr = f3(1.0f);
    const float cur_ior = 1.0f;
    uint32_t depth = 0;
    bool delta = true;
    F3 L = f3(0.0f);
    const bool nee = rp.use_nee && !s.lights.empty();
    while (true) {
        rc.ext++;
        const HitRec h = rp.tprim ? closest_tprim(s, ray) : closest_ref(s, ray, nullptr, nullptr);
        if (rc.log) {
            PathRecord pr{0.0f, {ray.o.x, ray.o.y, ray.o.z}, ray.tmin, {ray.d.x, ray.d.y, ray.d.z}, ray.tmax,
                          {0, h.t, h.u, h.v}, (float)depth, {0, 0}};
            const uint32_t prim = h.tri == kInvalid ? kInvalid : s.tris[h.tri].prim;
            std::memcpy(&pr.result[0], &prim, 4);
            rc.log->push_back(pr);
        }
        if (h.tri == kInvalid) break;
        const Triangle& tr = s.tris[h.tri];
        const Material& m = s.materials[tr.material];
        if (tr.model == Emissive) {   // MaterialEval.cu:882-912
            if (!rp.use_nee || delta) L = L + thr * f3(m.emission[0], m.emission[1], m.emission[2]);
            break;
        }
        // LoadResolvedHit + GetShadingFrame, MaterialEval.cu:54-115
        const float b0 = 1.0f - h.u - h.v, b1 = h.u, b2 = h.v;
        const F3 pos = ray.o + ray.d * h.t;
        const F3 gn = fix_normal(tr.gn, ray.d);
        const F3 nrm = b0 * tr.n[0] + b1 * tr.n[1] + b2 * tr.n[2];
        float tg[4];
        for (int c = 0; c < 4; c++) tg[c] = b0 * tr.t[0][c] + b1 * tr.t[1][c] + b2 * tr.t[2][c];
        const F3 bt = f3(tg[0], tg[1], tg[2]);
        const float sign = tg[3] < 0.0f ? -1.0f : 1.0f;
        const Frame base = build_frame(nrm, bt, sign);
        const F3 nts = normalize(f3(m.normal_ts[0], m.normal_ts[1], m.normal_ts[2]));
        const F3 sn = nts.x * base.t + nts.y * base.b + nts.z * base.n;
        const Frame fr = build_frame(sn, bt, sign);
        SampleRng rng(rp.seed, sample, 2u + 8u * (uint64_t)depth);
        const float u0 = rng.next(), u1 = rng.next(), u2 = rng.next(), u3 = rng.next();
        const float l0 = rng.next(), l1 = rng.next(), l2 = rng.next();

        auto shadow = [&](F3 lp, F3 radiance) {   // shadow ray + ClampRadiance + AddRadiance
            const F3 so = pos + gn * 1e-4f;
            const F3 sv = lp - so;
            const float dist = length(sv);
            const Ray sr{so, sv / dist, kMinT, dist - 1e-4f};
            rc.shadow++;
            const bool occluded = any_ref(s, sr);
            if (rc.log)
                rc.log->push_back(PathRecord{1.0f, {sr.o.x, sr.o.y, sr.o.z}, sr.tmin, {sr.d.x, sr.d.y, sr.d.z}, sr.tmax,
                                             {occluded ? 1.0f : 0.0f, radiance.x, radiance.y, radiance.z},
                                             (float)depth, {0, 0}});
            if (!occluded) {
                const float l2s = dot(radiance, radiance);
                if (l2s > 20.0f * 20.0f) radiance = radiance * (20.0f * rsqrt_host(l2s));
                L = L + radiance;
            }
        };
        struct LightHit {
            F3 p, gn, wi, emission;
            float d2, pdf;
        };
        auto sample_light = [&](float us, float x0, float x1) {   // MaterialEval.cu:171-192
            float pmf;
            const uint32_t li = alias_sample(s.bins.data(), (uint32_t)s.bins.size(), us, &pmf);
            const Light& lt = s.lights[li];
            const float sq = std::sqrt(x0);
            const float c1 = sq * (1.0f - x1), c2 = sq * x1;
            LightHit lh;
            lh.p = lt.v0 + c1 * lt.e1 + c2 * lt.e2;
            const F3 to = lh.p - pos;
            lh.d2 = dot(to, to);
            lh.wi = normalize(to);
            lh.gn = fix_normal(lt.gn, lh.wi);
            lh.emission = lt.emission;
            lh.pdf = pmf / lt.area;
            return lh;
        };

        bool scatter = false, new_delta = false;
        F3 wi = f3(0.0f), weight = f3(0.0f);
        if (tr.model == Diffuse) {   // MaterialEval.cu:231-350
            const F3 albedo = f3(m.color[0], m.color[1], m.color[2]);
            if (nee) {
                const LightHit lh = sample_light(l0, l1, l2);
                if (!(lh.d2 < 1e-6f)) {
                    const float ci = std::fmax(dot(fr.n, lh.wi), 0.0f);
                    const float gci = std::fmax(dot(gn, lh.wi), 0.0f);
                    if (!(ci <= 0.0f || gci <= 0.0f)) {
                        const float cl = std::fmax(dot(lh.gn, f3(0.0f) - lh.wi), 0.0f);
                        if (!(cl <= 0.0f)) {
                            const F3 brdf = albedo / kPi;
                            const float G = ci * cl / lh.d2;
                            shadow(lh.p, lh.emission * thr * brdf * G / lh.pdf);
                        }
                    }
                }
            }
            wi = fr.from_local(cosine_hemisphere(u0, u1));
            const float ci = std::fmax(dot(fr.n, wi), 0.0f);
            const float gci = std::fmax(dot(gn, wi), 0.0f);
            if (!(ci <= 0.0f || gci <= 0.0f)) {
                weight = albedo;
                scatter = true;
            }
        } else if (tr.model == Ggx) {   // MaterialEval.cu:652-880
            const GgxParams g = ggx_params(m, cur_ior);
            const F3 wo = fr.to_local(f3(0.0f) - ray.d);
            if (nee) {
                const LightHit lh = sample_light(l0, l1, l2);
                if (!(lh.d2 < 1e-6f || lh.pdf <= 0.0f)) {
                    const float gci = dot(gn, lh.wi);
                    if (!(gci <= 0.0f)) {
                        const float cl = std::fmax(dot(lh.gn, f3(0.0f) - lh.wi), 0.0f);
                        if (!(cl <= 0.0f)) {
                            const F3 wl = fr.to_local(lh.wi);
                            F3 f;
                            if (ggx_direct_bsdf(g, wo, wl, &f)) {
                                const float G = wl.z * cl / lh.d2;
                                shadow(lh.p, lh.emission * thr * f * G / lh.pdf);
                            }
                        }
                    }
                }
            }
            GgxSample gs;
            if (ggx_indirect(g, wo, u0, u1, u2, &gs)) {
                const F3 ww = fr.from_local(gs.wl);
                if (!(dot(ww, gn) <= 0.0f)) {
                    wi = ww;
                    weight = gs.weight;
                    scatter = true;
                }
            }
        } else if (tr.model == Mirror) {   // MaterialEval.cu:352-392
            wi = reflect(ray.d, fr.n);
            weight = f3(m.specular[0], m.specular[1], m.specular[2]);
            scatter = true;
            new_delta = true;
        } else {
            break;   // Normal / Phong are outside the core
        }
        if (!scatter) break;
        // ComputePathContinuation, MaterialEval.cu:129-146
        const uint32_t next = depth + 1;
        if (next >= rp.depth) break;
        F3 nt = thr * weight;
        if (next >= rp.rr_start) {
            const float q = clampf(max_component(nt), 0.0f, 1.0f);
            if (u3 > q) break;
            nt = nt / q;
        }
        // SpawnScatteredRay, MaterialEval.cu:148-156
        ray = Ray{ray.o + ray.d * h.t + gn * 1e-4f, wi, kMinT, kMaxT};
        thr = nt;
        depth = next;
        delta = new_delta;
    }
    return L;
}

template <class F>
void parallel_for(uint64_t n, uint32_t threads, F&& f) {
    if (threads <= 1 || n < 256) {
        for (uint64_t i = 0; i < n; i++) f(i);
        return;
    }
    std::atomic<uint64_t> next{0};
    std::vector<std::thread> pool;
    for (uint32_t k = 0; k < threads; k++)
        pool.emplace_back([&]() {
            while (true) {
                const uint64_t b = next.fetch_add(256);
                if (b >= n) break;
                const uint64_t e = std::min<uint64_t>(n, b + 256);
                for (uint64_t i = b; i < e; i++) f(i);
            }
        });
    for (auto& t : pool) t.join();
}

// ------------------------------------------------------------------ walk of the core's BVH4
// Mirrors the visit order of the GPU closest-hit kernel over the 128-byte BVH4
// nodes the core builds: slab test in the form t = lo*(1/d) - o*(1/d) with a
// fused multiply-add, the four children sorted nearest first by a 5-exchange
// network, farther children pushed with their entry distance, popped entries
// beyond the current hit skipped.  Hits are the (t, prim id) minimum, so they
// must equal closest_brute(); the visit counts are the algorithmic-byte model
// of the roofline.
struct Bvh4Walk {
    uint32_t nodes = 0, tris = 0, max_stack = 0;
};

inline float safe_rcp(float x) {
    const float ax = std::fabs(x) < 1e-20f ? std::copysign(1e-20f, x) : x;
    return 1.0f / ax;
}

// Scale 2^(e - 127) of a quantized node's axis, and byte k of a bound row.
inline float qscale(uint32_t e) {
    const uint32_t b = e << 23;
    float f;
    std::memcpy(&f, &b, 4);
    return f;
}
inline float qbyte(uint32_t row, int k) { return (float)((row >> (8 * k)) & 0xFFu); }

bool bvh4q_any(const Scene& s, const std::vector<uint32_t>& by_prim, const uint32_t* qnodes, const uint32_t* order,
               const Ray& ray, Bvh4Walk& w) {
    const F3 rd = f3(safe_rcp(ray.d.x), safe_rcp(ray.d.y), safe_rcp(ray.d.z));
    const F3 ord = ray.o * rd;
    std::vector<uint32_t> stack;
    uint32_t ref = 0;
    const float inf = std::numeric_limits<float>::infinity();
    while (true) {
        if (!(ref & 0x80000000u)) {
            // 16 words: origin x, y, z | scale exponents | lo x, lo y, lo z, hi x | hi y, hi z, child 0, 1 | child 2, 3
            const uint32_t* q = qnodes + 16 * (size_t)ref;
            w.nodes++;
            float o[3];
            std::memcpy(o, q, sizeof(o));
            const float ax = qscale(q[3] & 0xFFu) * rd.x, ay = qscale((q[3] >> 8) & 0xFFu) * rd.y,
                        az = qscale((q[3] >> 16) & 0xFFu) * rd.z;
            const float bx = std::fma(o[0], rd.x, -ord.x), by = std::fma(o[1], rd.y, -ord.y),
                        bz = std::fma(o[2], rd.z, -ord.z);
            // the near plane of every child is its lo row where the direction is >= 0
            const bool px = rd.x >= 0.0f, py = rd.y >= 0.0f, pz = rd.z >= 0.0f;
            const uint32_t nx = px ? q[4] : q[7], fx = px ? q[7] : q[4];
            const uint32_t ny = py ? q[5] : q[8], fy = py ? q[8] : q[5];
            const uint32_t nz = pz ? q[6] : q[9], fz = pz ? q[9] : q[6];
            uint32_t c[4] = {q[10], q[11], q[12], q[13]};
            float t0[4];
            for (int k = 0; k < 4; k++) {
                const float tnx = std::fma(qbyte(nx, k), ax, bx), tfx = std::fma(qbyte(fx, k), ax, bx);
                const float tny = std::fma(qbyte(ny, k), ay, by), tfy = std::fma(qbyte(fy, k), ay, by);
                const float tnz = std::fma(qbyte(nz, k), az, bz), tfz = std::fma(qbyte(fz, k), az, bz);
                const float tn = std::fmax(std::fmax(tnx, tny), std::fmax(tnz, ray.tmin));
                const float tf = std::fmin(std::fmin(tfx, tfy), std::fmin(tfz, ray.tmax));
                t0[k] = tn <= tf ? tn : inf;
            }
            auto cs = [&](int a, int b) {
                if (t0[b] < t0[a]) {
                    std::swap(t0[a], t0[b]);
                    std::swap(c[a], c[b]);
                }
            };
            cs(0, 1);
            cs(2, 3);
            cs(0, 2);
            cs(1, 3);
            cs(1, 2);
            uint32_t next = kInvalid;
            for (int k = 3; k >= 0; k--) {
                if (t0[k] != inf && c[k] != kInvalid) {
                    if (next != kInvalid) stack.push_back(next);
                    next = c[k];
                }
            }
            w.max_stack = std::max(w.max_stack, (uint32_t)stack.size());
            if (next != kInvalid) {
                ref = next;
                continue;
            }
        } else {
            const uint32_t first = (ref & 0x7FFFFFFFu) >> 4, cnt = (ref & 15u) + 1u;
            w.tris += cnt;
            bool hit = false;
            for (uint32_t k = 0; k < cnt; k++) {
                float t, u, v;
                if (mt(s.tris[by_prim[order[first + k]]], ray.o, ray.d, ray.tmin, t, u, v) && t < ray.tmax) hit = true;
            }
            if (hit) return true;
        }
        if (stack.empty()) return false;
        ref = stack.back();
        stack.pop_back();
    }
}

// Walks over the core's 8-wide compressed quantized nodes, in the kernel's
// order: a ray of octant oct (bit a set where its reciprocal direction
// component is negative) takes a node's entered children in visit order
// k ^ oct of their slots -- first every leaf child, then the inner children,
// the pending ones kept as one stack entry per node -- counting node steps and
// triangle tests.  ANY: the first hit ends the walk; otherwise the closest hit
// with the (t, prim) tie rule, every box tested against the current closest t.
// 32 words per node: origin x y z | scale exponents, inner mask in byte 3 | lo
// rows x0 x1 y0 y1 z0 z1 | hi rows | first inner child | first triangle |
// triangle ends of the slots (bytes) x2 | pad.  `order` maps the tree's
// triangle positions to prims.
template <bool ANY>
HitRec bvh8q_walk(const Scene& s, const std::vector<uint32_t>& by_prim, const uint32_t* qnodes, const uint32_t* order,
                  const Ray& ray, Bvh4Walk& w) {
    const F3 rd = f3(safe_rcp(ray.d.x), safe_rcp(ray.d.y), safe_rcp(ray.d.z));
    const F3 ord = ray.o * rd;
    const bool px = rd.x >= 0.0f, py = rd.y >= 0.0f, pz = rd.z >= 0.0f;
    const uint32_t oct = (px ? 0u : 1u) | (py ? 0u : 2u) | (pz ? 0u : 4u);
    float tmax = ray.tmax;
    HitRec best{kInvalid, 0, 0, 0};
    uint32_t best_prim = kInvalid;
    struct Group {
        uint32_t base, imask, pending;
    };
    std::vector<Group> stack;
    auto child = [&](uint32_t base, uint32_t imask, uint32_t pending) {
        const uint32_t slot = (uint32_t)__builtin_ctz(pending) ^ oct;
        return base + (uint32_t)__builtin_popcount(imask & ((1u << slot) - 1u));
    };
    uint32_t node = 0;
    while (true) {
        const uint32_t* q = qnodes + 32 * (size_t)node;
        w.nodes++;
        float o[3];
        std::memcpy(o, q, sizeof(o));
        const float ax = qscale(q[3] & 0xFFu) * rd.x, ay = qscale((q[3] >> 8) & 0xFFu) * rd.y,
                    az = qscale((q[3] >> 16) & 0xFFu) * rd.z;
        const float bx = std::fma(o[0], rd.x, -ord.x), by = std::fma(o[1], rd.y, -ord.y),
                    bz = std::fma(o[2], rd.z, -ord.z);
        const uint32_t* lo = q + 4;
        const uint32_t* hi = q + 10;
        const uint32_t imask = q[3] >> 24, node_base = q[16], tri_base = q[17];
        uint8_t end[8];
        for (int k = 0; k < 8; k++) end[k] = (uint8_t)(q[18 + (k >> 2)] >> (8 * (k & 3)));
        uint32_t inner_v = 0, leaf_v = 0;
        for (int k = 0; k < 8; k++) {
            const int wd = k >> 2, b = k & 3;
            const uint32_t nx = px ? lo[wd] : hi[wd], fx = px ? hi[wd] : lo[wd];
            const uint32_t ny = py ? lo[2 + wd] : hi[2 + wd], fy = py ? hi[2 + wd] : lo[2 + wd];
            const uint32_t nz = pz ? lo[4 + wd] : hi[4 + wd], fz = pz ? hi[4 + wd] : lo[4 + wd];
            const float tnx = std::fma(qbyte(nx, b), ax, bx), tfx = std::fma(qbyte(fx, b), ax, bx);
            const float tny = std::fma(qbyte(ny, b), ay, by), tfy = std::fma(qbyte(fy, b), ay, by);
            const float tnz = std::fma(qbyte(nz, b), az, bz), tfz = std::fma(qbyte(fz, b), az, bz);
            const float tn = std::fmax(std::fmax(tnx, tny), std::fmax(tnz, ray.tmin));
            const float tf = std::fmin(std::fmin(tfx, tfy), std::fmin(tfz, tmax));
            if (!(tn <= tf)) continue;
            const uint32_t pos = (uint32_t)k ^ oct;
            if (imask >> k & 1u) inner_v |= 1u << pos;
            else leaf_v |= 1u << pos;
        }
        uint32_t next = kInvalid;
        if (inner_v) {
            next = child(node_base, imask, inner_v);
            const uint32_t rest = inner_v & (inner_v - 1u);
            if (rest) stack.push_back(Group{node_base, imask, rest});
            w.max_stack = std::max(w.max_stack, (uint32_t)stack.size());
        }
        for (uint32_t lv = leaf_v; lv; lv &= lv - 1u) {
            const uint32_t slot = (uint32_t)__builtin_ctz(lv) ^ oct;
            const uint32_t first = slot ? end[slot - 1] : 0u, cnt = end[slot] - first;
            w.tris += cnt;
            bool hit = false;
            for (uint32_t k = 0; k < cnt; k++) {
                const uint32_t prim = order[tri_base + first + k];
                const uint32_t ti = by_prim[prim];
                float t, u, v;
                if (!mt(s.tris[ti], ray.o, ray.d, ray.tmin, t, u, v)) continue;
                if (ANY) {
                    if (t < tmax) {
                        hit = true;
                        best = HitRec{ti, t, u, v};
                    }
                } else if (t < tmax || (t == tmax && best_prim != kInvalid && prim < best_prim)) {
                    tmax = t;
                    best = HitRec{ti, t, u, v};
                    best_prim = prim;
                }
            }
            if (ANY && hit) return best;
        }
        if (next != kInvalid) {
            node = next;
            continue;
        }
        if (stack.empty()) return best;
        Group& g = stack.back();
        node = child(g.base, g.imask, g.pending);
        g.pending &= g.pending - 1u;
        if (!g.pending) stack.pop_back();
    }
}

bool bvh8q_any(const Scene& s, const std::vector<uint32_t>& by_prim, const uint32_t* qnodes, const uint32_t* order,
               const Ray& ray, Bvh4Walk& w) {
    return bvh8q_walk<true>(s, by_prim, qnodes, order, ray, w).tri != kInvalid;
}

HitRec bvh4_closest(const Scene& s, const std::vector<uint32_t>& by_prim, const uint32_t* nodes, const uint32_t* order,
                    const Ray& ray, Bvh4Walk& w) {
    const F3 rd = f3(safe_rcp(ray.d.x), safe_rcp(ray.d.y), safe_rcp(ray.d.z));
    const F3 ord = ray.o * rd;
    float tmax = ray.tmax;
    HitRec best{kInvalid, 0, 0, 0};
    uint32_t best_prim = kInvalid;
    std::vector<std::pair<uint32_t, float>> stack;
    uint32_t ref = 0;
    const float inf = std::numeric_limits<float>::infinity();
    while (true) {
        bool pop = false;
        if (!(ref & 0x80000000u)) {
            const uint32_t* nd = nodes + 32 * (size_t)ref;
            float f[24];
            std::memcpy(f, nd, sizeof(f));
            uint32_t c[4];
            std::memcpy(c, nd + 24, sizeof(c));
            w.nodes++;
            float t0[4];
            for (int i = 0; i < 4; i++) {
                const float ax = std::fma(f[0 + i], rd.x, -ord.x), bx = std::fma(f[4 + i], rd.x, -ord.x);
                const float ay = std::fma(f[8 + i], rd.y, -ord.y), by = std::fma(f[12 + i], rd.y, -ord.y);
                const float az = std::fma(f[16 + i], rd.z, -ord.z), bz = std::fma(f[20 + i], rd.z, -ord.z);
                const float tn = std::fmax(std::fmax(std::fmin(ax, bx), std::fmin(ay, by)), std::fmax(std::fmin(az, bz), ray.tmin));
                const float tf = std::fmin(std::fmin(std::fmax(ax, bx), std::fmax(ay, by)), std::fmin(std::fmax(az, bz), tmax));
                t0[i] = tn <= tf ? tn : inf;
            }
            auto cs = [&](int a, int b) {
                if (t0[b] < t0[a]) {
                    std::swap(t0[a], t0[b]);
                    std::swap(c[a], c[b]);
                }
            };
            cs(0, 1);
            cs(2, 3);
            cs(0, 2);
            cs(1, 3);
            cs(1, 2);
            uint32_t next = kInvalid;
            float tnext = 0.0f;
            for (int i = 3; i >= 0; i--) {
                if (t0[i] != inf && c[i] != kInvalid) {
                    if (next != kInvalid) stack.emplace_back(next, tnext);
                    next = c[i];
                    tnext = t0[i];
                }
            }
            w.max_stack = std::max(w.max_stack, (uint32_t)stack.size());
            if (next == kInvalid) pop = true;
            else ref = next;
        } else {
            const uint32_t first = (ref & 0x7FFFFFFFu) >> 4, cnt = (ref & 15u) + 1u;
            w.tris += cnt;
            for (uint32_t k = 0; k < cnt; k++) {
                const uint32_t prim = order[first + k];
                const uint32_t ti = by_prim[prim];
                float t, u, v;
                if (mt(s.tris[ti], ray.o, ray.d, ray.tmin, t, u, v)) {
                    if (t < tmax || (t == tmax && best_prim != kInvalid && prim < best_prim)) {
                        tmax = t;
                        best = HitRec{ti, t, u, v};
                        best_prim = prim;
                    }
                }
            }
            pop = true;
        }
        if (pop) {
            bool found = false;
            while (!stack.empty()) {
                const auto e = stack.back();
                stack.pop_back();
                if (e.second <= tmax) {
                    ref = e.first;
                    found = true;
                    break;
                }
            }
            if (!found) break;
        }
    }
    return best;
}

}  // namespace orc

// ====================================================================== C ABI for the tests
using namespace orc;

extern "C" {

void* orc_scene_create(uint32_t T, const float* positions, const float* normals, const float* tangents,
                       const uint32_t* material_ids, uint32_t nmat, const void* materials, const float* light_weights) {
    auto* s = new Scene();
    s->materials.resize(nmat);
    std::memcpy(s->materials.data(), materials, sizeof(Material) * nmat);
    s->tris.resize(T);
    std::vector<float> lw;
    std::vector<uint32_t> li;
    for (uint32_t t = 0; t < T; t++) {
        Triangle& tr = s->tris[t];
        const float* P = positions + 9 * (size_t)t;
        for (int k = 0; k < 3; k++) {
            tr.p[k] = f3(P[3 * k], P[3 * k + 1], P[3 * k + 2]);
            const float* N = normals + 9 * (size_t)t + 3 * k;
            tr.n[k] = f3(N[0], N[1], N[2]);
            if (tangents) {
                for (int c = 0; c < 4; c++) tr.t[k][c] = tangents[12 * (size_t)t + 4 * k + c];
            } else {
                const F3 axis = std::fabs(tr.n[k].z) < 0.999f ? f3(0.0f, 0.0f, 1.0f) : f3(1.0f, 0.0f, 0.0f);
                const F3 fb = normalize(cross(axis, tr.n[k]));
                tr.t[k][0] = fb.x;
                tr.t[k][1] = fb.y;
                tr.t[k][2] = fb.z;
                tr.t[k][3] = 1.0f;
            }
        }
        tr.e1 = tr.p[1] - tr.p[0];   // BuildTriangles, Renderer.cpp:200-207
        tr.e2 = tr.p[2] - tr.p[0];
        tr.v0 = tr.p[0];
        tr.material = material_ids[t];
        tr.model = s->materials[tr.material].model;
        tr.gn = normalize(cross(tr.e1, tr.e2));
        tr.prim = t;
        if (tr.model == Emissive) {   // LightSampler.cpp:5-40
            const float area = 0.5f * length(cross(tr.e1, tr.e2));
            lw.push_back(light_weights ? light_weights[t] : area);
            li.push_back((uint32_t)s->lights.size());
            const Material& m = s->materials[tr.material];
            s->lights.push_back(Light{tr.e1, tr.e2, tr.v0, t, tr.gn, area, f3(m.emission[0], m.emission[1], m.emission[2])});
        }
    }
    if (!s->lights.empty()) s->bins = build_alias(lw, li);
    s->max_depth = estimate_depth(T);
    s->min_triangles = estimate_leaf(T);
    if (T > 0) {
        auto root = build_recurse(*s, 0, T, 0);
        flatten(*root, s->nodes);
    }
    s->by_prim.resize(T);
    for (uint32_t i = 0; i < T; i++) s->by_prim[s->tris[i].prim] = i;
    return s;
}

void orc_scene_destroy(void* h) { delete static_cast<Scene*>(h); }

void orc_scene_info(void* h, uint32_t* out6) {
    Scene* s = static_cast<Scene*>(h);
    out6[0] = (uint32_t)s->nodes.size();
    out6[1] = s->depth;
    out6[2] = s->max_depth;
    out6[3] = s->min_triangles;
    out6[4] = (uint32_t)s->lights.size();
    out6[5] = (uint32_t)s->tris.size();
}

// nodes as 10 u32/floats each (min xyz, max xyz, left, right, first, count); prims: BVH order -> prim
void orc_scene_bvh(void* h, float* nodes_out, uint32_t* prim_order) {
    Scene* s = static_cast<Scene*>(h);
    for (size_t i = 0; i < s->nodes.size(); i++) {
        const Node& n = s->nodes[i];
        float* o = nodes_out + 10 * i;
        o[0] = n.bounds.min.x; o[1] = n.bounds.min.y; o[2] = n.bounds.min.z;
        o[3] = n.bounds.max.x; o[4] = n.bounds.max.y; o[5] = n.bounds.max.z;
        std::memcpy(o + 6, &n.left, 4);
        std::memcpy(o + 7, &n.right, 4);
        std::memcpy(o + 8, &n.first, 4);
        std::memcpy(o + 9, &n.count, 4);
    }
    for (size_t i = 0; i < s->tris.size(); i++) prim_order[i] = s->tris[i].prim;
}

// rays: n x 8 floats (o, tmin, d, tmax); hits: n x (prim u32, t, u, v)
// mode 0 = reference traversal (first found), 1 = brute force (t, prim) minimum
void orc_trace_closest(void* h, const float* rays, uint32_t n, void* hits, int mode, uint32_t threads,
                       uint32_t* nodes_visited, uint32_t* tris_tested) {
    const Scene& s = *static_cast<Scene*>(h);
    parallel_for(n, threads, [&](uint64_t i) {
        const float* r = rays + 8 * i;
        const Ray ray{f3(r[0], r[1], r[2]), f3(r[4], r[5], r[6]), r[3], r[7]};
        HitRec hr;
        if (s.tris.empty()) hr = HitRec{kInvalid, 0, 0, 0};
        else if (mode == 0)
            hr = closest_ref(s, ray, nodes_visited ? nodes_visited + i : nullptr, tris_tested ? tris_tested + i : nullptr);
        else if (mode == 2) hr = closest_tprim(s, ray);
        else hr = closest_brute(s, ray);
        uint32_t* o = static_cast<uint32_t*>(hits) + 4 * i;
        if (hr.tri == kInvalid) {
            o[0] = kInvalid;
            float z = 0.0f;
            std::memcpy(o + 1, &z, 4);
            std::memcpy(o + 2, &z, 4);
            std::memcpy(o + 3, &z, 4);
        } else {
            o[0] = s.tris[hr.tri].prim;
            std::memcpy(o + 1, &hr.t, 4);
            std::memcpy(o + 2, &hr.u, 4);
            std::memcpy(o + 3, &hr.v, 4);
        }
    });
}

void orc_trace_any(void* h, const float* rays, uint32_t n, uint32_t* occluded, int mode, uint32_t threads) {
    const Scene& s = *static_cast<Scene*>(h);
    parallel_for(n, threads, [&](uint64_t i) {
        const float* r = rays + 8 * i;
        const Ray ray{f3(r[0], r[1], r[2]), f3(r[4], r[5], r[6]), r[3], r[7]};
        occluded[i] = s.tris.empty() ? 0u : (mode == 0 ? any_ref(s, ray) : any_brute(s, ray)) ? 1u : 0u;
    });
}

// hit -> material queue (0 diffuse, 1 ggx, 2 mirror, 3 emissive) or 4 regen, Intersect.cu:115-122
void orc_classify(void* h, const float* rays, uint32_t n, uint32_t* queue, uint32_t threads) {
    const Scene& s = *static_cast<Scene*>(h);
    parallel_for(n, threads, [&](uint64_t i) {
        const float* r = rays + 8 * i;
        const Ray ray{f3(r[0], r[1], r[2]), f3(r[4], r[5], r[6]), r[3], r[7]};
        const HitRec hr = s.tris.empty() ? HitRec{kInvalid, 0, 0, 0} : closest_tprim(s, ray);
        uint32_t q = 4;
        if (hr.tri != kInvalid) {
            switch (s.tris[hr.tri].model) {
                case Diffuse: q = 0; break;
                case Ggx: q = 1; break;
                case Mirror: q = 2; break;
                case Emissive: q = 3; break;
                default: q = 4; break;
            }
        }
        queue[i] = q;
    });
}

// Closest hit over the core's BVH4 (nodes: N x 32 u32, order: BVH slot -> prim).
// Any hit over the core's 64-byte quantized BVH4 in its kernel's visit order
// (nearest child first, the farther ones pushed; a hit ends the ray after its
// leaf), counting node steps and triangle tests per ray: the visits behind the
// any-hit traversal's algorithmic bytes.
void orc_bvh4q_any(void* h, const uint32_t* qnodes, const uint32_t* order, const float* rays, uint32_t n,
                   uint32_t* occluded, uint32_t* nodes_visited, uint32_t* tris_tested, uint32_t threads) {
    const Scene& s = *static_cast<Scene*>(h);
    parallel_for(n, threads, [&](uint64_t i) {
        const float* r = rays + 8 * i;
        const Ray ray{f3(r[0], r[1], r[2]), f3(r[4], r[5], r[6]), r[3], r[7]};
        Bvh4Walk w;
        occluded[i] = bvh4q_any(s, s.by_prim, qnodes, order, ray, w) ? 1u : 0u;
        if (nodes_visited) nodes_visited[i] = w.nodes;
        if (tris_tested) tris_tested[i] = w.tris;
    });
}

void orc_bvh8q_any(void* h, const uint32_t* qnodes, const uint32_t* order, const float* rays, uint32_t n,
                   uint32_t* occluded, uint32_t* nodes_visited, uint32_t* tris_tested, uint32_t* max_stack,
                   uint32_t threads) {
    const Scene& s = *static_cast<Scene*>(h);
    parallel_for(n, threads, [&](uint64_t i) {
        const float* r = rays + 8 * i;
        const Ray ray{f3(r[0], r[1], r[2]), f3(r[4], r[5], r[6]), r[3], r[7]};
        Bvh4Walk w;
        occluded[i] = bvh8q_any(s, s.by_prim, qnodes, order, ray, w) ? 1u : 0u;
        if (nodes_visited) nodes_visited[i] = w.nodes;
        if (tris_tested) tris_tested[i] = w.tris;
        if (max_stack) max_stack[i] = w.max_stack;
    });
}

void orc_bvh8q_closest(void* h, const uint32_t* qnodes, const uint32_t* order, const float* rays, uint32_t n,
                       void* hits, uint32_t* nodes_visited, uint32_t* tris_tested, uint32_t* max_stack,
                       uint32_t threads) {
    const Scene& s = *static_cast<Scene*>(h);
    parallel_for(n, threads, [&](uint64_t i) {
        const float* r = rays + 8 * i;
        const Ray ray{f3(r[0], r[1], r[2]), f3(r[4], r[5], r[6]), r[3], r[7]};
        Bvh4Walk w;
        const HitRec hr = bvh8q_walk<false>(s, s.by_prim, qnodes, order, ray, w);
        uint32_t* o = static_cast<uint32_t*>(hits) + 4 * i;
        o[0] = hr.tri == kInvalid ? kInvalid : s.tris[hr.tri].prim;
        const float z = 0.0f;
        std::memcpy(o + 1, hr.tri == kInvalid ? &z : &hr.t, 4);
        std::memcpy(o + 2, hr.tri == kInvalid ? &z : &hr.u, 4);
        std::memcpy(o + 3, hr.tri == kInvalid ? &z : &hr.v, 4);
        if (nodes_visited) nodes_visited[i] = w.nodes;
        if (tris_tested) tris_tested[i] = w.tris;
        if (max_stack) max_stack[i] = w.max_stack;
    });
}

void orc_bvh4_closest(void* h, const uint32_t* nodes, const uint32_t* order, const float* rays, uint32_t n, void* hits,
                      uint32_t* nodes_visited, uint32_t* tris_tested, uint32_t* max_stack, uint32_t threads) {
    const Scene& s = *static_cast<Scene*>(h);
    parallel_for(n, threads, [&](uint64_t i) {
        const float* r = rays + 8 * i;
        const Ray ray{f3(r[0], r[1], r[2]), f3(r[4], r[5], r[6]), r[3], r[7]};
        Bvh4Walk w;
        const HitRec hr = bvh4_closest(s, s.by_prim, nodes, order, ray, w);
        uint32_t* o = static_cast<uint32_t*>(hits) + 4 * i;
        o[0] = hr.tri == kInvalid ? kInvalid : s.tris[hr.tri].prim;
        const float z = 0.0f;
        std::memcpy(o + 1, hr.tri == kInvalid ? &z : &hr.t, 4);
        std::memcpy(o + 2, hr.tri == kInvalid ? &z : &hr.u, 4);
        std::memcpy(o + 3, hr.tri == kInvalid ? &z : &hr.v, 4);
        if (nodes_visited) nodes_visited[i] = w.nodes;
        if (tris_tested) tris_tested[i] = w.tris;
        if (max_stack) max_stack[i] = w.max_stack;
    });
}

void orc_alias_build(const float* w, uint32_t n, void* bins_out) {
    std::vector<float> wv(w, w + n);
    std::vector<uint32_t> vals(n);
    for (uint32_t i = 0; i < n; i++) vals[i] = i;
    const std::vector<Bin> b = build_alias(wv, vals);
    std::memcpy(bins_out, b.data(), sizeof(Bin) * n);
}

void orc_alias_sample(const void* bins, uint32_t size, const float* u, uint32_t n, uint32_t* idx, float* pmf) {
    const Bin* b = static_cast<const Bin*>(bins);
    for (uint32_t i = 0; i < n; i++) idx[i] = alias_sample(b, size, u[i], pmf + i);
}

void orc_scene_alias(void* h, void* bins_out) {
    const Scene& s = *static_cast<Scene*>(h);
    std::memcpy(bins_out, s.bins.data(), sizeof(Bin) * s.bins.size());
}

void orc_pcg32(uint64_t initstate, uint64_t initseq, uint64_t advance, uint32_t n, uint32_t* out) {
    Pcg r = pcg_srandom(initstate, initseq);
    pcg_advance(r, advance);
    for (uint32_t i = 0; i < n; i++) out[i] = pcg_next(r);
}

// MakeDeviceCamera, Renderer.cpp:237-257; quaternion (x, y, z, w) rotation as glm does it.
void orc_camera_basis(float fov_y, float aspect, const float* pos, const float* q, float* out12) {
    auto rot = [&](F3 v) {
        const F3 qv = f3(q[0], q[1], q[2]);
        const F3 uv = cross(qv, v);
        const F3 uuv = cross(qv, uv);
        return v + ((uv * q[3]) + uuv) * 2.0f;
    };
    const F3 fwd = rot(f3(0, 0, -1)), right = rot(f3(1, 0, 0)), up = rot(f3(0, 1, 0));
    const float halfH = std::tan(fov_y * 0.5f);
    const float halfW = halfH * aspect;
    const F3 origin = f3(pos[0], pos[1], pos[2]);
    const F3 hor = 2.0f * halfW * right;
    const F3 ver = 2.0f * halfH * up;
    const F3 llc = origin + fwd - halfW * right - halfH * up;
    const F3 v4[4] = {origin, llc, hor, ver};
    for (int i = 0; i < 4; i++) {
        out12[3 * i] = v4[i].x;
        out12[3 * i + 1] = v4[i].y;
        out12[3 * i + 2] = v4[i].z;
    }
}

// Renders pixels [pixel_begin, pixel_end) of a W x H frame; hdr receives
// accumulated radiance / spp as RGBA float (alpha 0) for those pixels, in
// double-precision accumulation.  settings: width, height, spp, depth, use_nee,
// rr_start; camera12 from orc_camera_basis.
static RenderParams make_params(const uint32_t* settings, uint64_t seed, const float* camera12) {
    RenderParams rp;
    rp.width = settings[0];
    rp.height = settings[1];
    const uint32_t spp_in = std::max(1u, settings[2]);
    uint32_t axis = (uint32_t)std::sqrt((double)spp_in);
    while ((uint64_t)(axis + 1) * (axis + 1) <= spp_in) axis++;
    while ((uint64_t)axis * axis > spp_in) axis--;
    rp.axis = axis;
    rp.spp = axis * axis;
    rp.depth = std::min(std::max(settings[3], 1u), 100u);
    rp.use_nee = settings[4];
    rp.rr_start = settings[5] ? settings[5] : 3u;
    rp.tprim = settings[6];
    rp.seed = seed;
    rp.origin = f3(camera12[0], camera12[1], camera12[2]);
    rp.llc = f3(camera12[3], camera12[4], camera12[5]);
    rp.horizontal = f3(camera12[6], camera12[7], camera12[8]);
    rp.vertical = f3(camera12[9], camera12[10], camera12[11]);
    return rp;
}

void orc_render(void* h, const uint32_t* settings, uint64_t seed, const float* camera12, uint32_t pixel_begin,
                uint32_t pixel_end, uint32_t threads, float* hdr, uint64_t* counters2) {
    const Scene& s = *static_cast<Scene*>(h);
    const RenderParams rp = make_params(settings, seed, camera12);
    const uint32_t npix = pixel_end - pixel_begin;
    std::atomic<uint64_t> ext{0}, sh{0};
    parallel_for(npix, threads, [&](uint64_t k) {
        const uint32_t pixel = pixel_begin + (uint32_t)k;
        double acc[3] = {0, 0, 0};
        RenderCounters rc;
        for (uint32_t c = 0; c < rp.spp; c++) {
            const uint64_t sample = (uint64_t)pixel * rp.spp + c;
            uint32_t px;
            const F3 L = trace_sample(s, rp, sample, rc, &px);
            acc[0] += L.x;
            acc[1] += L.y;
            acc[2] += L.z;
        }
        float* o = hdr + 4 * (size_t)k;
        o[0] = (float)(acc[0] / rp.spp);
        o[1] = (float)(acc[1] / rp.spp);
        o[2] = (float)(acc[2] / rp.spp);
        o[3] = 0.0f;
        ext.fetch_add(rc.ext, std::memory_order_relaxed);
        sh.fetch_add(rc.shadow, std::memory_order_relaxed);
    });
    if (counters2) {
        counters2[0] = ext.load();
        counters2[1] = sh.load();
    }
}

// The path of one sample, as every traced ray with its result (16 floats per
// record, PathRecord), for comparing a sample's rays with the core's
// traversal hooks; returns the record count (records beyond `capacity` are
// dropped) and the sample's radiance in L3.
uint32_t orc_trace_path(void* h, const uint32_t* settings, uint64_t seed, const float* camera12, uint64_t sample,
                        float* records, uint32_t capacity, float* L3) {
    const Scene& s = *static_cast<Scene*>(h);
    RenderParams rp = make_params(settings, seed, camera12);
    std::vector<PathRecord> log;
    RenderCounters rc;
    rc.log = &log;
    uint32_t px;
    const F3 L = trace_sample(s, rp, sample, rc, &px);
    L3[0] = L.x, L3[1] = L.y, L3[2] = L.z;
    const uint32_t n = (uint32_t)std::min<size_t>(log.size(), capacity);
    std::memcpy(records, log.data(), n * sizeof(PathRecord));
    return (uint32_t)log.size();
}

// PostprocessKernel, Image.cu:34-54: hdr already divided by spp.
void orc_tonemap(const float* hdr, uint32_t n, uint32_t* rgba) {
    const float inv_gamma = 1.0f / 2.2f;
    for (uint32_t i = 0; i < n; i++) {
        uint32_t out = 0xFF000000u;
        for (int k = 0; k < 3; k++) {
            const float c = hdr[4 * (size_t)i + k];
            const float t = c / (c + 1.0f);
            const float g = std::pow(t, inv_gamma);
            const uint32_t b = (uint32_t)(unsigned char)clampf(g * 255.0f, 0.0f, 255.0f);
            out |= b << (8 * k);
        }
        rgba[i] = out;
    }
}


// ---- shading functions, exported for the closed-form tests (test_oracle.py)

void orc_sincos(const float* x, uint32_t n, float* s, float* c) {
    for (uint32_t i = 0; i < n; i++) sincos_angle(x[i], s + i, c + i);
}

void orc_cosine_hemisphere(const float* u2, uint32_t n, float* w3) {
    for (uint32_t i = 0; i < n; i++) {
        const F3 w = cosine_hemisphere(u2[2 * i], u2[2 * i + 1]);
        w3[3 * i] = w.x, w3[3 * i + 1] = w.y, w3[3 * i + 2] = w.z;
    }
}

static F3 ld3(const float* p, uint32_t i) { return f3(p[3 * i], p[3 * i + 1], p[3 * i + 2]); }
static void st3(float* p, uint32_t i, F3 v) { p[3 * i] = v.x, p[3 * i + 1] = v.y, p[3 * i + 2] = v.z; }

// D(wm) (what = 0), Lambda(wm) (1), G1(wm) (2) of one alpha
void orc_ggx_term(int what, float alpha, const float* wm3, uint32_t n, float* out) {
    for (uint32_t i = 0; i < n; i++) {
        const F3 w = ld3(wm3, i);
        out[i] = what == 0 ? dggx(w, alpha) : what == 1 ? lambda(w, alpha) : g1(w, alpha);
    }
}

// D_wo(wm) = G1(wo) / |wo.z| D(wm) |wo.wm| (the visible-normal density)
void orc_ggx_dvis(float alpha, const float* wo3, const float* wm3, uint32_t n, float* out) {
    for (uint32_t i = 0; i < n; i++) out[i] = dvis(ld3(wo3, i), ld3(wm3, i), alpha);
}

void orc_ggx_sample_wm(float alpha, const float* wo3, const float* u2, uint32_t n, float* wm3) {
    for (uint32_t i = 0; i < n; i++) st3(wm3, i, sample_wm(ld3(wo3, i), alpha, u2[2 * i], u2[2 * i + 1]));
}

// material = {r, g, b, roughness, metallic, ior}; the incoming medium has ior 1
static GgxParams params_of(const float* m6) {
    Material m{};
    m.model = Ggx;
    m.color[0] = m6[0], m.color[1] = m6[1], m.color[2] = m6[2];
    m.roughness = m6[3];
    m.metallic = m6[4];
    m.ior = m6[5];
    return ggx_params(m, 1.0f);
}

void orc_ggx_bsdf(const float* m6, const float* wo3, const float* wl3, uint32_t n, float* f3out, uint32_t* ok) {
    const GgxParams g = params_of(m6);
    for (uint32_t i = 0; i < n; i++) {
        F3 f = f3(0.0f);
        ok[i] = ggx_direct_bsdf(g, ld3(wo3, i), ld3(wl3, i), &f) ? 1u : 0u;
        st3(f3out, i, f);
    }
}

// density of the indirect sample's direction at wl (z > 0 and wo.wm > 0), 0 otherwise
void orc_ggx_pdf(const float* m6, const float* wo3, const float* wl3, uint32_t n, float* pdf) {
    const GgxParams g = params_of(m6);
    float pd = 0.0f, ps = 0.0f;
    const bool lobes = ggx_lobes(g, &pd, &ps);
    for (uint32_t i = 0; i < n; i++) {
        const F3 wo = ld3(wo3, i), wl = ld3(wl3, i);
        pdf[i] = 0.0f;
        if (!lobes || wl.z <= 0.0f) continue;
        const F3 wm = normalize(wl + wo);
        const float owm = dot(wo, wm);
        if (owm <= 0.0f) continue;
        pdf[i] = ggx_pdf(pd, ps, wo, wl, wm, owm, g.alpha);
    }
}

void orc_ggx_sample(const float* m6, const float* wo3, const float* u3, uint32_t n, float* wl3, float* f3out,
                    float* weight3, float* pdf, uint32_t* ok) {
    const GgxParams g = params_of(m6);
    for (uint32_t i = 0; i < n; i++) {
        GgxSample gs{f3(0.0f), f3(0.0f), f3(0.0f), 0.0f};
        ok[i] = ggx_indirect(g, ld3(wo3, i), u3[3 * i], u3[3 * i + 1], u3[3 * i + 2], &gs) ? 1u : 0u;
        st3(wl3, i, gs.wl);
        st3(f3out, i, gs.f);
        st3(weight3, i, gs.weight);
        pdf[i] = gs.pdf;
    }
}

}  // extern "C"
