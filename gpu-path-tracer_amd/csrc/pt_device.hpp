// Shared host/device definitions of the wavefront core: record layouts in HBM,
// float3 arithmetic, the counter-addressable PCG32 sample streams, and the
// per-sample math (camera, frames, BSDFs) that the kernels run.
//
// Every translation unit is compiled with -ffp-contract=off and IEEE division
// and square root, so a float expression written here rounds the same way on
// the CPU and on gfx950.  That is what makes hit ids and queue membership
// reproducible bit for bit by an independent CPU implementation.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define PT_HD __host__ __device__ __forceinline__

namespace pt {

constexpr uint32_t kInvalid = 0xFFFFFFFFu;
constexpr float kMinT = 1e-4f;          // PathTracerDefaults.hpp:13
constexpr float kMaxT = 1e34f;          // PathTracerDefaults.hpp:14
constexpr float kPi = 3.14159265358979323846f;
constexpr float kInvPi = 1.0f / 3.14159265358979323846f;
constexpr float kOneMinusEpsilon = 0x1.fffffep-1f;   // Math.hpp:11
constexpr uint32_t kMaxDepthLimit = 100;             // PathTracerDefaults.hpp:18
constexpr uint32_t kQueueCount = 4;                  // diffuse, ggx, mirror, emissive

// ---------------------------------------------------------------- float3 math
struct V3 {
    float x, y, z;
};

PT_HD V3 v3(float x, float y, float z) { return V3{x, y, z}; }
PT_HD V3 v3(float s) { return V3{s, s, s}; }
PT_HD V3 operator+(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
PT_HD V3 operator-(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
PT_HD V3 operator-(V3 a) { return V3{-a.x, -a.y, -a.z}; }
PT_HD V3 operator*(V3 a, V3 b) { return V3{a.x * b.x, a.y * b.y, a.z * b.z}; }
PT_HD V3 operator*(V3 a, float s) { return V3{a.x * s, a.y * s, a.z * s}; }
PT_HD V3 operator*(float s, V3 a) { return V3{s * a.x, s * a.y, s * a.z}; }
PT_HD V3 operator/(V3 a, float s) { return V3{a.x / s, a.y / s, a.z / s}; }
PT_HD V3 operator/(V3 a, V3 b) { return V3{a.x / b.x, a.y / b.y, a.z / b.z}; }
PT_HD float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
PT_HD V3 cross(V3 a, V3 b) {
    return V3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
PT_HD float length(V3 a) { return sqrtf(dot(a, a)); }
// One definition of normalize on both sides: v * (1/sqrt(v.v)), IEEE ops.
PT_HD V3 normalize(V3 a) { return a * (1.0f / sqrtf(dot(a, a))); }
PT_HD V3 reflect(V3 i, V3 n) { return i - 2.0f * n * dot(n, i); }
PT_HD float maxc(V3 a) { return fmaxf(a.x, fmaxf(a.y, a.z)); }
PT_HD float clampf(float v, float lo, float hi) { return fminf(fmaxf(v, lo), hi); }
PT_HD V3 lerp3(V3 a, V3 b, float t) { return a + t * (b - a); }
PT_HD float sqr(float v) { return v * v; }

PT_HD V3 xyz(float4 a) { return V3{a.x, a.y, a.z}; }
PT_HD float4 f4(V3 a, float w) { return make_float4(a.x, a.y, a.z, w); }

// ---------------------------------------------------------------- records in HBM

// BVH4 node: one 128-byte line.  Six 4-wide coordinate rows so one lane's box
// test is 24 independent slab terms, then four child references.  A reference
// with bit 31 clear is a node index; with bit 31 set it is a leaf:
// bits 4..30 = first triangle (BVH order), bits 0..3 = count - 1.
struct alignas(128) Bvh4Node {
    float lox[4], hix[4], loy[4], hiy[4], loz[4], hiz[4];
    uint32_t child[4];
    uint32_t info[4];   // per-child triangle count (leaf) or 0; diagnostics only
};
static_assert(sizeof(Bvh4Node) == 128, "BVH4 node must fill one cache line");

// BVH2 node: 64 bytes, half a line, the same references and empty-slot rule.
// Rows are (lo0, lo1, hi0, hi1) per axis, so one float4 load serves one axis.
struct alignas(64) Bvh2Node {
    float lox[2], hix[2], loy[2], hiy[2], loz[2], hiz[2];
    uint32_t child[2];
    uint32_t info[2];
};
static_assert(sizeof(Bvh2Node) == 64, "BVH2 node is 64 bytes");

// Quantized BVH4 node: 64 bytes, the same four references, child boxes as
// 8-bit offsets from the node's own box corner.  Axis a of child i decodes to
//   lo = origin[a] + float(lo byte i of row a) * 2^(e_a - 127)
// (likewise hi), in float arithmetic without contraction; the quantizer picks
// each byte so that the decoded box contains the child's fp32 box bit for bit,
// so every ray that enters the fp32 box enters the decoded one.
struct alignas(64) Bvh4QNode {
    float origin[3];
    uint32_t scale;      // bytes 0..2: float exponent field e_a of axis a's scale
    uint32_t lo[3];      // row per axis, byte i = child i
    uint32_t hi[3];
    uint32_t child[4];
    uint32_t pad[2];
};
static_assert(sizeof(Bvh4QNode) == 64, "quantized BVH4 node is 64 bytes");

PT_HD float qnode_decode(float origin, uint32_t q, float scale) { return origin + (float)q * scale; }
PT_HD float qnode_scale(uint32_t e) { return __builtin_bit_cast(float, e << 23); }

// Smallest scale exponent (1..254) whose byte 255 reaches `hi` from `origin`.
PT_HD uint32_t qnode_exponent(float origin, float hi) {
    auto reaches = [&](uint32_t e) { return qnode_decode(origin, 255u, qnode_scale(e)) >= hi; };
    const float extent = hi - origin;
    uint32_t e = 1;
    if (extent > 0.0f) {
        // 2^(e - 127) >= extent / 255 as a first guess, then the exact minimum
        const uint32_t ebits = (__builtin_bit_cast(uint32_t, extent * (1.0f / 255.0f)) >> 23) & 0xFFu;
        e = ebits < 1u ? 1u : (ebits > 253u ? 254u : ebits + 1u);
        while (e > 1u && reaches(e - 1u)) e--;
        while (e < 254u && !reaches(e)) e++;
    }
    return e;
}

// Largest byte whose decoded value is <= lo, and smallest whose value is >= hi.
PT_HD uint32_t qnode_floor(float origin, float scale, float lo) {
    uint32_t q = 0;
    for (uint32_t step = 128; step > 0; step >>= 1)
        if (q + step <= 255u && qnode_decode(origin, q + step, scale) <= lo) q += step;
    return q;
}
PT_HD uint32_t qnode_ceil(float origin, float scale, float hi) {
    uint32_t q = 255;
    for (uint32_t step = 128; step > 0; step >>= 1)
        if (q >= step && qnode_decode(origin, q - step, scale) >= hi) q -= step;
    return q;
}

// fp32 BVH4 node -> quantized node (host and device give identical bytes).
// Empty child slots (reference 0xFFFFFFFF) keep lo 255, hi 0.
PT_HD void quantize_bvh4_node(const Bvh4Node& n, Bvh4QNode& q) {
    const float* lo[3] = {n.lox, n.loy, n.loz};
    const float* hi[3] = {n.hix, n.hiy, n.hiz};
    q.scale = 0u;
    q.pad[0] = q.pad[1] = 0u;
    for (int a = 0; a < 3; a++) {
        float mn = 0.0f, mx = 0.0f;
        bool any = false;
        for (int i = 0; i < 4; i++) {
            if (n.child[i] == 0xFFFFFFFFu) continue;
            mn = any ? (lo[a][i] < mn ? lo[a][i] : mn) : lo[a][i];
            mx = any ? (hi[a][i] > mx ? hi[a][i] : mx) : hi[a][i];
            any = true;
        }
        const uint32_t e = qnode_exponent(mn, mx);
        const float s = qnode_scale(e);
        q.origin[a] = mn;
        q.scale |= e << (8 * a);
        uint32_t rl = 0, rh = 0;
        for (int i = 0; i < 4; i++) {
            uint32_t bl = 255u, bh = 0u;
            if (n.child[i] != 0xFFFFFFFFu) {
                bl = qnode_floor(mn, s, lo[a][i]);
                bh = qnode_ceil(mn, s, hi[a][i]);
            }
            rl |= bl << (8 * i);
            rh |= bh << (8 * i);
        }
        q.lo[a] = rl;
        q.hi[a] = rh;
    }
    for (int i = 0; i < 4; i++) q.child[i] = n.child[i];
}

// Quantized BVH8 node of the any-hit tree, in the manner of the compressed
// wide BVH of Ylitie, Karras and Laine (HPG 2017): eight children under the
// byte scheme of Bvh4QNode, bound rows holding children 0-3 in word 2a and 4-7
// in word 2a + 1 of axis a.  Children are implicit: the inner children of a
// node are consecutive nodes from node_base in slot order (breadth-first
// numbering), the triangles of its leaf children consecutive from tri_base in
// slot order, and byte s of tri_end the end of slot s's triangles relative to
// tri_base (a prefix sum: the start of slot s is byte s - 1, 0 for slot 0).
// Byte 3 of `scale` marks the inner slots.  The node is read as five 16-byte
// loads: (origin, scale), (lo x0 x1 y0 y1), (lo z0 z1, hi x0 x1),
// (hi y0 y1 z0 z1), (node_base, tri_base, tri_end); the line's last 48 bytes
// are padding.
struct alignas(128) Bvh8QNode {
    float origin[3];
    uint32_t scale;
    uint32_t lo[6];
    uint32_t hi[6];
    uint32_t node_base;
    uint32_t tri_base;
    uint32_t tri_end[2];
    uint32_t pad[12];
};
static_assert(sizeof(Bvh8QNode) == 128, "quantized BVH8 node fills one cache line");

// Quantizes `n` child boxes (lo/hi [axis][child], padded fp32; ref 0xFFFFFFFF
// = empty slot) into origin, scale exponents and byte rows (byte i of row word
// a * words + i / 4); the rule of quantize_bvh4_node.
PT_HD void quantize_children(const float lo[3][8], const float hi[3][8], const uint32_t* child, int n,
                             float origin[3], uint32_t& scale, uint32_t* lo_rows, uint32_t* hi_rows) {
    const int words = (n + 3) / 4;
    scale = 0u;
    for (int a = 0; a < 3; a++) {
        float mn = 0.0f, mx = 0.0f;
        bool any = false;
        for (int i = 0; i < n; i++) {
            if (child[i] == 0xFFFFFFFFu) continue;
            mn = any ? (lo[a][i] < mn ? lo[a][i] : mn) : lo[a][i];
            mx = any ? (hi[a][i] > mx ? hi[a][i] : mx) : hi[a][i];
            any = true;
        }
        const uint32_t e = qnode_exponent(mn, mx);
        const float s = qnode_scale(e);
        origin[a] = mn;
        scale |= e << (8 * a);
        for (int w = 0; w < words; w++) {
            uint32_t rl = 0, rh = 0;
            for (int k = 0; k < 4; k++) {
                const int i = 4 * w + k;
                uint32_t bl = 255u, bh = 0u;
                if (i < n && child[i] != 0xFFFFFFFFu) {
                    bl = qnode_floor(mn, s, lo[a][i]);
                    bh = qnode_ceil(mn, s, hi[a][i]);
                }
                rl |= bl << (8 * k);
                rh |= bh << (8 * k);
            }
            lo_rows[a * words + w] = rl;
            hi_rows[a * words + w] = rh;
        }
    }
}

constexpr uint32_t kLeafBit = 0x80000000u;
PT_HD bool ref_is_leaf(uint32_t r) { return (r & kLeafBit) != 0u; }
PT_HD uint32_t leaf_first(uint32_t r) { return (r & 0x7FFFFFFFu) >> 4; }
PT_HD uint32_t leaf_count(uint32_t r) { return (r & 15u) + 1u; }
PT_HD uint32_t make_leaf_ref(uint32_t first, uint32_t count) {
    return kLeafBit | (first << 4) | (count - 1u);
}
constexpr uint32_t kMaxLeafSize = 16;
constexpr uint32_t kMaxLeafFirst = (1u << 27) - 1u;

// Top-of-tree table: the first kTopBytes / sizeof(node) nodes of a tree in
// breadth-first order, copied with every internal child reference that lands
// in the table rewritten to kTopBit | its table index, followed by the global
// index of each table node (uint32 each).  Traversal blocks keep the table in
// LDS; an internal reference with kTopBit set reads its node there.  Internal
// node indices stay below kTopBit.
constexpr uint32_t kTopBit = 0x40000000u;
constexpr uint32_t kTopBytes = 8192;
constexpr uint32_t kMaxNodeIndex = kTopBit - 1u;
PT_HD bool ref_is_top(uint32_t r) { return (r & (kLeafBit | kTopBit)) == kTopBit; }

// Intersection record, 48 B, in BVH order (TriangleIntersection, Triangle.hpp:8-17):
//   v0.xyz | original prim id,  e1.xyz | queue id,  e2.xyz | material id
struct alignas(16) TriRecord {
    float v0[3];
    uint32_t prim;
    float e1[3];
    uint32_t queue;
    float e2[3];
    uint32_t material;
};
static_assert(sizeof(TriRecord) == 48, "triangle record is 48 B");

// Shading record, 112 B, by original prim id (TriangleShading, Triangle.hpp:19-31,
// without the UVs: every material here is constant).
struct alignas(16) ShadeRecord {
    float n[3][4];      // corner normals (w unused)
    float t[3][4];      // corner tangents, w = handedness
    float gn[3];        // geometric normal normalize(cross(e1, e2))
    uint32_t material;
};
static_assert(sizeof(ShadeRecord) == 112, "shading record is 112 B");

// Constant material, 80 B.
struct alignas(16) MaterialRecord {
    float color[3];
    uint32_t queue;
    float specular[3];
    float alpha;        // max(roughness^2, 1e-3), MaterialEval.cu:860
    float metallic;
    float ior;
    uint32_t model;
    float pad0;
    float emission[3];
    float pad1;
    float nts[3];       // normalize(normal_ts), the decoded normal-map constant
    float pad2;
};
static_assert(sizeof(MaterialRecord) == 80, "material record is 80 B");

// Emissive triangle, 80 B (LightTriangle, Triangle.hpp:40-58, constant emission).
struct alignas(16) LightRecord {
    float v0[3];
    float area;
    float e1[3];
    uint32_t prim;      // original prim id (LightTriangle.index)
    float e2[3];
    float pad0;
    float gn[3];
    float pad1;
    float emission[3];
    float pad2;
};
static_assert(sizeof(LightRecord) == 80, "light record is 80 B");

struct AliasBin {
    float q, p;
    uint32_t alias;
    uint32_t value;
};

// Per-path flags in meta.y: bits 0..7 depth, bits 8..23 view id, bit 31 = last
// scatter was delta.
constexpr uint32_t kDeltaBit = 0x80000000u;
PT_HD uint32_t flags_depth(uint32_t f) { return f & 0xFFu; }
PT_HD uint32_t flags_view(uint32_t f) { return (f >> 8) & 0xFFFFu; }

// ---------------------------------------------------------------- PCG32
// PCG-XSH-RR 64/32 (O'Neill 2014, pcg-random.org).  A sample owns the stream
// `sample index`; draw n of that sample is the n-th output of
// pcg32_srandom(seed, sample).  Jump-ahead is Brown's arbitrary-stride LCG
// advance (1994), so any (sample, draw) pair is addressable without state.
constexpr uint64_t kPcgMult = 6364136223846793005ull;

PT_HD uint32_t pcg_output(uint64_t old) {
    uint32_t xorshifted = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    uint32_t rot = (uint32_t)(old >> 59u);
    return (xorshifted >> rot) | (xorshifted << ((0u - rot) & 31u));
}

struct PcgStream {
    uint64_t state;
    uint64_t inc;
};

PT_HD PcgStream pcg_seed(uint64_t initstate, uint64_t initseq) {
    PcgStream s;
    s.state = 0u;
    s.inc = (initseq << 1u) | 1u;
    s.state = s.state * kPcgMult + s.inc;
    s.state += initstate;
    s.state = s.state * kPcgMult + s.inc;
    return s;
}

// Affine map of an n-step advance with unit increment: state' = mult*state + plus*inc.
struct PcgJump {
    uint64_t mult;
    uint64_t plus;
};

PT_HD PcgJump pcg_jump(uint64_t delta) {
    uint64_t cur_mult = kPcgMult, cur_plus = 1u, acc_mult = 1u, acc_plus = 0u;
    while (delta > 0u) {
        if (delta & 1u) {
            acc_mult *= cur_mult;
            acc_plus = acc_plus * cur_mult + cur_plus;
        }
        cur_plus = (cur_mult + 1u) * cur_plus;
        cur_mult *= cur_mult;
        delta >>= 1u;
    }
    return PcgJump{acc_mult, acc_plus};
}

PT_HD float u01(uint32_t x) { return (float)(x >> 8) * 0x1p-24f; }

// Draw layout of one sample: draws 0,1 are the camera jitter; bounce b owns
// draws 2+8b .. 9+8b: slot 0,1 direction, 2 lobe, 3 Russian roulette,
// 4 light pick, 5,6 point on the light, 7 spare.
constexpr uint32_t kDrawsPerBounce = 8;
PT_HD uint64_t bounce_draw_base(uint32_t bounce) { return 2u + (uint64_t)kDrawsPerBounce * bounce; }

// ---------------------------------------------------------------- camera
struct CameraBasis {           // DeviceCamera, DeviceCamera.hpp:5-11
    float origin[3];
    float llc[3];
    float horizontal[3];
    float vertical[3];
};

// One camera view of one resident scene: a frame of the single-frame API, or
// one element of a batch or of the serving stream.  Views live in a device
// ring of kViewRing records addressed by a monotonically growing virtual
// index; the ring slot (virtual & (kViewRing-1)) is what a path carries.  A
// record may be one chunk of a view: it owns the global samples
// [sample_begin, sample_begin + sample_count), which are the view's local
// samples [sample_offset, sample_offset + sample_count), and the view's pixels
// [accum_offset, accum_offset + w*h) of the accumulation buffer.  Sample
// streams are keyed by (seed, local sample index), so a view renders the same
// alone, in a batch or interleaved with other requests.
struct ViewRecord {
    CameraBasis cam;
    uint32_t width, height, spp, spp_axis;
    uint32_t depth_limit, use_nee, rr_start, scene;
    uint64_t seed;
    uint64_t sample_begin;
    uint64_t sample_count;
    uint64_t sample_offset;
    uint32_t accum_offset;
    float inv_spp;
};
static_assert(sizeof(ViewRecord) == 120, "view record is 120 B");
constexpr uint32_t kViewRing = 4096;       // ring slot lives in bits 8..23 of the path flags
constexpr uint32_t kMaxScenes = 64;

// Wavefront constants in device memory, read by every kernel of the bounce
// loop so that one captured graph serves every frame.  Views with virtual
// index in [view_lo, view_hi) are live; views[view_lo] starts at or before the
// next sample to launch, so the generator finds a sample's view by binary
// search over sample_begin in that window.  total_samples is the global end of
// the samples admitted so far (it grows while a server runs).
struct FrameParams {
    uint64_t total_samples;
    uint32_t view_lo, view_hi;
    PcgJump bounce_jump[kMaxDepthLimit + 1];   // advance to draw 2+8b
};

// Counters of the wavefront, one block of device memory.
struct Counters {
    uint32_t ray[2];
    uint32_t queue[kQueueCount + 1];   // diffuse, ggx, mirror, emissive, regen
    uint32_t shadow[2];         // shadow rays pushed by the shade of an iteration of parity p
    uint32_t work_closest;
    uint32_t work_shadow[2];
    uint32_t active;            // ray count of the next iteration
    uint32_t iterations;
    uint32_t invalid_pixels;
    uint32_t stack_overflow;
    uint32_t records;           // guide records of this run: record slots [0, records) are written
    unsigned long long launched;
    unsigned long long ext_rays;
    unsigned long long shadow_rays;
    unsigned long long closest_nodes;
    unsigned long long closest_tris;
    unsigned long long shadow_nodes;
    unsigned long long shadow_tris;
    unsigned long long dg_entries;   // diffuse and GGX queue entries shaded by this run
};

}  // namespace pt
