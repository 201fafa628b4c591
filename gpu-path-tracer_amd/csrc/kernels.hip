// Wavefront path-tracing kernels for gfx950 (CDNA4, wave64).
//
// Stages of one bounce (the loop of Renderer::RenderFrame,
// Core/Source/Graphics/Cuda/PathTracing/Renderer.cpp:517-623):
//   closest  - persistent BVH4 traversal with per-lane ray refill from
//              per-wave chunks and an LDS stack; writes each entry's hit and
//              target queue by queue position
//   bin      - scatters queue positions into the diffuse / ggx / mirror /
//              emissive queues (five atomics per block)
//   shade    - one launch over the four material queues; continuing paths are
//              written compacted into the next ray queue, NEE shadow rays
//              into the shadow queue
//   shadow   - persistent any-hit traversal of the shadow queue
//   generate - camera paths for the free tail of the next ray queue
//   prepare  - one thread: counters, next ray count, ray statistics
// Queue space is reserved per block or per wave (ballot + mbcnt), never per
// lane: same-address atomics are expensive across the eight XCDs.
#include <hip/hip_runtime.h>

#include "guide_net.hpp"
#include "launch.hpp"

#include <mutex>

namespace pt {

// ================================================================ wave helpers

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

__device__ __forceinline__ uint32_t mbcnt(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

__device__ __forceinline__ uint32_t readlane_u32(uint32_t v, uint32_t lane) {
    return (uint32_t)__builtin_amdgcn_readlane((int)v, (int)lane);
}

// Every active lane calls; lanes with `pred` receive consecutive positions
// reserved with one atomicAdd on `counter` for the whole wave.
__device__ __forceinline__ uint32_t wave_reserve(unsigned int* counter, bool pred) {
    const uint64_t mask = __ballot(pred);
    uint32_t pos = 0;
    if (mask) {
        const uint32_t leader = (uint32_t)__ffsll((unsigned long long)mask) - 1u;
        uint32_t base = 0;
        if (lane_id() == leader) base = atomicAdd(counter, (unsigned int)__popcll(mask));
        base = readlane_u32(base, leader);
        pos = base + mbcnt(mask);
    }
    return pos;
}

// Multi-queue variant: `q` in [0, nq) or -1.  One ballot + atomic per distinct
// queue present in the wave.
__device__ __forceinline__ uint32_t wave_reserve_multi(unsigned int* counters, int q) {
    uint32_t pos = 0;
    uint64_t pending = __ballot(q >= 0);
    while (pending) {
        const uint32_t leader = (uint32_t)__ffsll((unsigned long long)pending) - 1u;
        const int lq = (int)readlane_u32((uint32_t)q, leader);
        const uint64_t m = __ballot(q == lq);
        uint32_t base = 0;
        if (lane_id() == leader) base = atomicAdd(counters + lq, (unsigned int)__popcll(m));
        base = readlane_u32(base, leader);
        if (q == lq) pos = base + mbcnt(m);
        pending &= ~m;
    }
    return pos;
}

__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

// Adds per-lane instrumentation totals with one atomic per wave.
__device__ __forceinline__ void add_counts(unsigned long long* nodes_ctr, unsigned long long* tris_ctr, uint32_t nodes,
                                           uint32_t tris) {
    const uint32_t n = wave_sum(nodes), t = wave_sum(tris);
    if (lane_id() == 0) {
        atomicAdd(nodes_ctr, (unsigned long long)n);
        atomicAdd(tris_ctr, (unsigned long long)t);
    }
}

// ================================================================ intersection

#ifndef PT_MT_BRANCHFREE
#define PT_MT_BRANCHFREE 1
#endif
#ifndef PT_ANY_SORTED
#define PT_ANY_SORTED 1    // any-hit rays over 4-wide nodes also visit the nearest child first
                           // (occluders near the origin end the ray sooner: -14% shadow time
                           // measured); the 8-wide tree orders children by octant slots
#endif
#ifndef PT_TRACE_WAVES
#define PT_TRACE_WAVES 6   // minimum waves per SIMD of the traversal kernels: 80 VGPRs, which every
#endif                     // one-scene kernel fits anyway and the several-scene closest-hit kernel
                           // reaches without spilling (82 otherwise: 5 waves); 7 or 8 measured slower
#ifndef PT_TOP_TABLE
#define PT_TOP_TABLE 1     // single-scene traversal blocks read the top of the tree from LDS
#endif
#ifndef PT_WHILE_WHILE
#define PT_WHILE_WHILE 0   // if-if measured faster on gfx950 (one step per outer iteration)
#endif

// Moller-Trumbore with the reference's operation order and tests
// (Intersect.cu:25-61).  The values of accepted hits are the same float
// operations as the reference's.  Returns true if t lies in [tmin, +inf); the
// caller applies its own upper bound.  The branch-free form evaluates every
// term and forms one accept mask at the end; the other exits at the first
// failed test as the reference does.
__device__ __forceinline__ bool intersect_triangle(const float4 a, const float4 b, const float4 c, V3 o, V3 d,
                                                   float tmin, float& t, float& u, float& v) {
    const V3 v0 = xyz(a), e1 = xyz(b), e2 = xyz(c);
    const V3 pv = cross(d, e2);
    const float det = dot(e1, pv);
#if PT_MT_BRANCHFREE
    const float inv_det = 1.0f / det;
    const V3 s = o - v0;
    u = dot(s, pv) * inv_det;
    const V3 qv = cross(s, e1);
    v = dot(d, qv) * inv_det;
    t = dot(e2, qv) * inv_det;
    const bool ok_det = !(fabsf(det) < 1e-7f);
    const bool ok_u = !(u < 0.0f || u > 1.0f);
    const bool ok_v = !(v < 0.0f || u + v > 1.0f);
    return ok_det & ok_u & ok_v & (t >= tmin);
#else
    if (fabsf(det) < 1e-7f) return false;
    const float inv_det = 1.0f / det;
    const V3 s = o - v0;
    u = dot(s, pv) * inv_det;
    if (u < 0.0f || u > 1.0f) return false;
    const V3 qv = cross(s, e1);
    v = dot(d, qv) * inv_det;
    if (v < 0.0f || u + v > 1.0f) return false;
    t = dot(e2, qv) * inv_det;
    return t >= tmin;
#endif
}

__device__ __forceinline__ void load_tri(const TriRecord* tri, float4& a, float4& b, float4& c) {
    const float4* p = reinterpret_cast<const float4*>(tri);
    a = p[0];
    b = p[1];
    c = p[2];
}

struct RayFrame {
    V3 o, d;
    V3 rd;      // 1/d with zero components replaced by +-1e-20
    V3 ord;     // o * rd
    float tmin, tmax;
};

__device__ __forceinline__ float safe_rcp(float x) {
    const float ax = fabsf(x) < 1e-20f ? copysignf(1e-20f, x) : x;
    return 1.0f / ax;
}

__device__ __forceinline__ void setup_ray(RayFrame& r, V3 o, V3 d, float tmin, float tmax) {
    r.o = o;
    r.d = d;
    r.rd = v3(safe_rcp(d.x), safe_rcp(d.y), safe_rcp(d.z));
    r.ord = o * r.rd;
    r.tmin = tmin;
    r.tmax = tmax;
}

// Entry distance of the four children of a node, +inf for a miss.
__device__ __forceinline__ void box4(const Bvh4Node* node, const RayFrame& r, float tmax, float t0[4]) {
    const float4* p = reinterpret_cast<const float4*>(node);
    const float4 lx = p[0], hx = p[1], ly = p[2], hy = p[3], lz = p[4], hz = p[5];
    const float lxa[4] = {lx.x, lx.y, lx.z, lx.w}, hxa[4] = {hx.x, hx.y, hx.z, hx.w};
    const float lya[4] = {ly.x, ly.y, ly.z, ly.w}, hya[4] = {hy.x, hy.y, hy.z, hy.w};
    const float lza[4] = {lz.x, lz.y, lz.z, lz.w}, hza[4] = {hz.x, hz.y, hz.z, hz.w};
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const float ax = fmaf(lxa[i], r.rd.x, -r.ord.x), bx = fmaf(hxa[i], r.rd.x, -r.ord.x);
        const float ay = fmaf(lya[i], r.rd.y, -r.ord.y), by = fmaf(hya[i], r.rd.y, -r.ord.y);
        const float az = fmaf(lza[i], r.rd.z, -r.ord.z), bz = fmaf(hza[i], r.rd.z, -r.ord.z);
        const float tn = fmaxf(fmaxf(fminf(ax, bx), fminf(ay, by)), fmaxf(fminf(az, bz), r.tmin));
        const float tf = fminf(fminf(fmaxf(ax, bx), fmaxf(ay, by)), fminf(fmaxf(az, bz), tmax));
        t0[i] = tn <= tf ? tn : __builtin_inff();
    }
}

// `arity` children per node of each node kind.
template <int K>
struct NodeOf;
template <>
struct NodeOf<kNodeBvh4> {
    using type = Bvh4Node;
    static constexpr int arity = 4;
};
template <>
struct NodeOf<kNodeBvh2> {
    using type = Bvh2Node;
    static constexpr int arity = 2;
};
template <>
struct NodeOf<kNodeBvh4Q> {
    using type = Bvh4QNode;
    static constexpr int arity = 4;
};

// Byte i of a quantized row as a float (one v_cvt_f32_ubyte<i>).
template <int I>
__device__ __forceinline__ float qbyte(uint32_t row) {
    return (float)((row >> (8 * I)) & 0xFFu);
}

// Entry distance of child I of a quantized node: per axis the plane slab
// distance q*A + B (A = scale*rd, B = origin*rd - o*rd) of the near and far
// byte rows, then the usual interval intersection with [tmin, tmax].
template <int I>
__device__ __forceinline__ float qchild(uint32_t nx, uint32_t fx, uint32_t ny, uint32_t fy, uint32_t nz, uint32_t fz,
                                        float ax, float bx, float ay, float by, float az, float bz,
                                        const RayFrame& r) {
    const float tnx = fmaf(qbyte<I>(nx), ax, bx), tfx = fmaf(qbyte<I>(fx), ax, bx);
    const float tny = fmaf(qbyte<I>(ny), ay, by), tfy = fmaf(qbyte<I>(fy), ay, by);
    const float tnz = fmaf(qbyte<I>(nz), az, bz), tfz = fmaf(qbyte<I>(fz), az, bz);
    const float tn = fmaxf(fmaxf(tnx, tny), fmaxf(tnz, r.tmin));
    const float tf = fminf(fminf(tfx, tfy), fminf(tfz, r.tmax));
    return tn <= tf ? tn : __builtin_inff();
}

// Quantized node: the four child boxes of one 64-byte line (origin and scale
// exponents, lo rows, hi rows, references) under the slab test.  A plane q of
// axis a lies at origin_a + q*scale_a, the product exact (scale_a is a power
// of two), so its slab distance (origin_a + q*scale_a - o_a)*rd_a is formed as
// q*(scale_a*rd_a) + (origin_a*rd_a - o_a*rd_a): scale_a*rd_a is again exact
// and each distance takes one fma and one byte conversion.  The distance is
// monotonic in q with the sign of rd_a, so the near plane of every child is
// the lo row where rd_a >= 0 and the hi row otherwise, and no per-child
// min/max is needed.  An empty slot (lo byte 255, hi byte 0) has its near
// plane beyond its far plane; its reference is kInvalid as well.
__device__ __forceinline__ void children(const Bvh4QNode* node, const RayFrame& r, float t0[4], uint32_t c[4]) {
    const uint4* p = reinterpret_cast<const uint4*>(node);
    const uint4 a = p[0], b = p[1], d = p[2], e = p[3];
    const float ax = qnode_scale(a.w & 0xFFu) * r.rd.x;
    const float ay = qnode_scale((a.w >> 8) & 0xFFu) * r.rd.y;
    const float az = qnode_scale((a.w >> 16) & 0xFFu) * r.rd.z;
    const float bx = fmaf(__uint_as_float(a.x), r.rd.x, -r.ord.x);
    const float by = fmaf(__uint_as_float(a.y), r.rd.y, -r.ord.y);
    const float bz = fmaf(__uint_as_float(a.z), r.rd.z, -r.ord.z);
    // rows: lo x, y, z = b.x, b.y, b.z; hi x, y, z = b.w, d.x, d.y
    const bool px = r.rd.x >= 0.0f, py = r.rd.y >= 0.0f, pz = r.rd.z >= 0.0f;
    const uint32_t nx = px ? b.x : b.w, fx = px ? b.w : b.x;
    const uint32_t ny = py ? b.y : d.x, fy = py ? d.x : b.y;
    const uint32_t nz = pz ? b.z : d.y, fz = pz ? d.y : b.z;
    t0[0] = qchild<0>(nx, fx, ny, fy, nz, fz, ax, bx, ay, by, az, bz, r);
    t0[1] = qchild<1>(nx, fx, ny, fy, nz, fz, ax, bx, ay, by, az, bz, r);
    t0[2] = qchild<2>(nx, fx, ny, fy, nz, fz, ax, bx, ay, by, az, bz, r);
    t0[3] = qchild<3>(nx, fx, ny, fy, nz, fz, ax, bx, ay, by, az, bz, r);
    c[0] = d.z;
    c[1] = d.w;
    c[2] = e.x;
    c[3] = e.y;
}

// Compressed 8-wide node (Bvh8QNode): the slab test of qchild over its eight
// children, as a mask of the children the segment enters (bit s = slot s; an
// empty slot's near plane lies beyond its far plane), with the node's child
// bookkeeping: inner-slot mask, first inner child, first triangle and the
// triangle ends of the slots.
struct Node8Hits {
    uint32_t hits, imask, node_base, tri_base, end0, end1;
};
template <int I>
__device__ __forceinline__ bool qenter(uint32_t nx, uint32_t fx, uint32_t ny, uint32_t fy, uint32_t nz, uint32_t fz,
                                       float ax, float bx, float ay, float by, float az, float bz, const RayFrame& r) {
    const float tnx = fmaf(qbyte<I>(nx), ax, bx), tfx = fmaf(qbyte<I>(fx), ax, bx);
    const float tny = fmaf(qbyte<I>(ny), ay, by), tfy = fmaf(qbyte<I>(fy), ay, by);
    const float tnz = fmaf(qbyte<I>(nz), az, bz), tfz = fmaf(qbyte<I>(fz), az, bz);
    const float tn = fmaxf(fmaxf(tnx, tny), fmaxf(tnz, r.tmin));
    const float tf = fminf(fminf(tfx, tfy), fminf(tfz, r.tmax));
    return tn <= tf;
}

__device__ __forceinline__ Node8Hits enter8(const Bvh8QNode* node, const RayFrame& r) {
    const uint4* p = reinterpret_cast<const uint4*>(node);
    const uint4 a = p[0], l0 = p[1], l1 = p[2], h1 = p[3], m = p[4];
    const float ax = qnode_scale(a.w & 0xFFu) * r.rd.x;
    const float ay = qnode_scale((a.w >> 8) & 0xFFu) * r.rd.y;
    const float az = qnode_scale((a.w >> 16) & 0xFFu) * r.rd.z;
    const float bx = fmaf(__uint_as_float(a.x), r.rd.x, -r.ord.x);
    const float by = fmaf(__uint_as_float(a.y), r.rd.y, -r.ord.y);
    const float bz = fmaf(__uint_as_float(a.z), r.rd.z, -r.ord.z);
    // rows: lo x = (l0.x, l0.y), lo y = (l0.z, l0.w), lo z = (l1.x, l1.y);
    //       hi x = (l1.z, l1.w), hi y = (h1.x, h1.y), hi z = (h1.z, h1.w)
    const bool px = r.rd.x >= 0.0f, py = r.rd.y >= 0.0f, pz = r.rd.z >= 0.0f;
    const uint32_t nx0 = px ? l0.x : l1.z, fx0 = px ? l1.z : l0.x, nx1 = px ? l0.y : l1.w, fx1 = px ? l1.w : l0.y;
    const uint32_t ny0 = py ? l0.z : h1.x, fy0 = py ? h1.x : l0.z, ny1 = py ? l0.w : h1.y, fy1 = py ? h1.y : l0.w;
    const uint32_t nz0 = pz ? l1.x : h1.z, fz0 = pz ? h1.z : l1.x, nz1 = pz ? l1.y : h1.w, fz1 = pz ? h1.w : l1.y;
    uint32_t h = 0u;
    h |= qenter<0>(nx0, fx0, ny0, fy0, nz0, fz0, ax, bx, ay, by, az, bz, r) ? 0x01u : 0u;
    h |= qenter<1>(nx0, fx0, ny0, fy0, nz0, fz0, ax, bx, ay, by, az, bz, r) ? 0x02u : 0u;
    h |= qenter<2>(nx0, fx0, ny0, fy0, nz0, fz0, ax, bx, ay, by, az, bz, r) ? 0x04u : 0u;
    h |= qenter<3>(nx0, fx0, ny0, fy0, nz0, fz0, ax, bx, ay, by, az, bz, r) ? 0x08u : 0u;
    h |= qenter<0>(nx1, fx1, ny1, fy1, nz1, fz1, ax, bx, ay, by, az, bz, r) ? 0x10u : 0u;
    h |= qenter<1>(nx1, fx1, ny1, fy1, nz1, fz1, ax, bx, ay, by, az, bz, r) ? 0x20u : 0u;
    h |= qenter<2>(nx1, fx1, ny1, fy1, nz1, fz1, ax, bx, ay, by, az, bz, r) ? 0x40u : 0u;
    h |= qenter<3>(nx1, fx1, ny1, fy1, nz1, fz1, ax, bx, ay, by, az, bz, r) ? 0x80u : 0u;
    return Node8Hits{h, a.w >> 24, m.x, m.y, m.z, m.w};
}

// Bits k and k ^ oct of the two bytes of m exchanged: slot order to the visit
// order of a ray of octant oct, and back.
__device__ __forceinline__ uint32_t octant_order(uint32_t m, uint32_t oct) {
    if (oct & 1u) m = ((m & 0x5555u) << 1) | ((m >> 1) & 0x5555u);
    if (oct & 2u) m = ((m & 0x3333u) << 2) | ((m >> 2) & 0x3333u);
    if (oct & 4u) m = ((m & 0x0F0Fu) << 4) | ((m >> 4) & 0x0F0Fu);
    return m;
}

__device__ __forceinline__ void children(const Bvh4Node* node, const RayFrame& r, float t0[4], uint32_t c[4]) {
    box4(node, r, r.tmax, t0);
    const uint4 cref = reinterpret_cast<const uint4*>(node)[6];
    c[0] = cref.x;
    c[1] = cref.y;
    c[2] = cref.z;
    c[3] = cref.w;
}

__device__ __forceinline__ void children(const Bvh2Node* node, const RayFrame& r, float t0[2], uint32_t c[2]) {
    const float4* p = reinterpret_cast<const float4*>(node);
    const float4 x = p[0], y = p[1], z = p[2];
    const uint4 cref = reinterpret_cast<const uint4*>(node)[3];
    const float lx[2] = {x.x, x.y}, hx[2] = {x.z, x.w};
    const float ly[2] = {y.x, y.y}, hy[2] = {y.z, y.w};
    const float lz[2] = {z.x, z.y}, hz[2] = {z.z, z.w};
#pragma unroll
    for (int i = 0; i < 2; i++) {
        const float ax = fmaf(lx[i], r.rd.x, -r.ord.x), bx = fmaf(hx[i], r.rd.x, -r.ord.x);
        const float ay = fmaf(ly[i], r.rd.y, -r.ord.y), by = fmaf(hy[i], r.rd.y, -r.ord.y);
        const float az = fmaf(lz[i], r.rd.z, -r.ord.z), bz = fmaf(hz[i], r.rd.z, -r.ord.z);
        const float tn = fmaxf(fmaxf(fminf(ax, bx), fminf(ay, by)), fmaxf(fminf(az, bz), r.tmin));
        const float tf = fminf(fminf(fmaxf(ax, bx), fmaxf(ay, by)), fminf(fmaxf(az, bz), r.tmax));
        t0[i] = tn <= tf ? tn : __builtin_inff();
    }
    c[0] = cref.x;
    c[1] = cref.y;
}

__device__ __forceinline__ void cswap(float& ta, uint32_t& ra, float& tb, uint32_t& rb) {
    if (tb < ta) {
        const float t = ta;
        ta = tb;
        tb = t;
        const uint32_t x = ra;
        ra = rb;
        rb = x;
    }
}

// ================================================================ persistent traversal
//
// One lane = one ray at a time.  When at least `refill` lanes of a wave are
// idle, the wave claims that many rays with one atomic and the idle lanes start
// on them, so the SIMD stays busy while long rays finish (dynamic fetch).
// The stack lives in LDS as [entry][lane] (bank-conflict free) and spills to a
// global area past `LDS` entries.  Each entry is (node ref, entry distance) so
// popped subtrees that lie beyond the current hit are skipped without a fetch.

template <int BLOCK, int LDS, int NODE, bool ANY, bool COUNT, bool MULTI, class Src>
__device__ __forceinline__ void trace_persistent(const Src& src, const SceneDev& sc, const SceneDev* scenes,
                                                 const TraceConfig& cfg) {
    using Node = typename NodeOf<NODE>::type;
    __shared__ uint2 stack[LDS][BLOCK];
    // The top of the tree in LDS (single-scene launches): the whole table from
    // 256 threads up, 32 bytes per thread below, so that small blocks keep
    // their occupancy; references past a block's share go back to HBM.
    constexpr uint32_t kTopMax = kTopBytes / sizeof(Node);
    constexpr uint32_t kTopCap =
        (MULTI || !PT_TOP_TABLE) ? 0u : (BLOCK >= 256 ? kTopBytes : (uint32_t)BLOCK * 32u) / (uint32_t)sizeof(Node);
    __shared__ Node top[kTopCap ? kTopCap : 1];
    const uint32_t tid = threadIdx.x;
    const void* top_src = sc.top[NODE == kNodeBvh4Q ? 1 : 0];
    const uint32_t root = (kTopCap && top_src) ? kTopBit : 0u;
    if (kTopCap && top_src) {
        const uint4* s4 = static_cast<const uint4*>(top_src);
        uint4* d4 = reinterpret_cast<uint4*>(top);
        constexpr uint32_t kWords = kTopCap * (uint32_t)(sizeof(Node) / 16);
        for (uint32_t i = tid; i < kWords; i += BLOCK) d4[i] = s4[i];
        __syncthreads();
        if (kTopCap < kTopMax) {
            const uint32_t* gidx =
                reinterpret_cast<const uint32_t*>(static_cast<const char*>(top_src) + kTopMax * sizeof(Node));
            constexpr uint32_t A = (uint32_t)NodeOf<NODE>::arity;
            for (uint32_t i = tid; i < kTopCap * A; i += BLOCK) {
                uint32_t& c = top[i / A].child[i % A];
                if (ref_is_top(c) && (c & (kTopBit - 1u)) >= kTopCap) c = gidx[c & (kTopBit - 1u)];
            }
            __syncthreads();
        }
    }
    const uint32_t gtid = blockIdx.x * BLOCK + tid;
    const uint32_t stride = gridDim.x * BLOCK;
    const uint32_t cap = min(cfg.stack_cap, (uint32_t)LDS + cfg.spill_depth);
    const uint32_t n = src.count();
    unsigned int* work = src.work();
    // chunk size from the ray count: about four chunks per wave of the grid,
    // one ray per lane at least, at most cfg.chunk
    const uint32_t waves = stride >> 6;
    const uint32_t chunk = min(max((n / (4u * waves) + 63u) & ~63u, 64u), max(cfg.chunk, 64u));

    bool active = false;      // lane holds a ray
    bool fin = false;         // its traversal has ended
    // Wave-uniform range of ray indices owned by this wave.  The first chunk of
    // every wave is assigned statically (no atomic: a launch would otherwise
    // start with one same-address atomic per wave); later chunks come from the
    // counter, offset past the statically assigned range.
    const uint32_t wave_id = gtid >> 6;
    const uint32_t static_end = waves * chunk;
    uint32_t chunk_next = min(wave_id * chunk, n), chunk_end = min(wave_id * chunk + chunk, n);
    bool exhausted = chunk_next >= n;
    uint32_t item = 0, ref = 0, sp = 0;
    // per lane when MULTI: the ray's scene
    const Node* nodes_base = static_cast<const Node*>(NODE == kNodeBvh4Q ? sc.qnodes : sc.nodes);
    const TriRecord* tris_base = sc.tris;
    RayFrame r;
    float best_t = 0.0f, best_u = 0.0f, best_v = 0.0f;
    uint32_t best_prim = kInvalid, best_queue = 0;
    uint32_t nodes = 0, tris = 0, acc_nodes = 0, acc_tris = 0;

    // next stack entry whose subtree can still hold a closer hit; fin when empty
    auto pop = [&]() {
        while (sp > 0) {
            sp--;
            const uint2 e = sp < (uint32_t)LDS ? stack[sp][tid] : cfg.spill[(size_t)(sp - LDS) * stride + gtid];
            if (ANY || __uint_as_float(e.y) <= r.tmax) {
                ref = e.x;
                return;
            }
        }
        fin = true;
    };
    auto push = [&](uint32_t node_ref, float t) {
        if (sp < cap) {
            const uint2 e = make_uint2(node_ref, __float_as_uint(t));
            if (sp < (uint32_t)LDS) stack[sp][tid] = e;
            else cfg.spill[(size_t)(sp - LDS) * stride + gtid] = e;
            sp++;
        } else {
            atomicOr(cfg.overflow, 1u);
        }
    };
    auto test_tri = [&](const float4& a, const float4& b, const float4& c) {
        float t, u, v;
        if (intersect_triangle(a, b, c, r.o, r.d, r.tmin, t, u, v)) {
            const uint32_t prim = __float_as_uint(a.w);
            if (ANY) {
                if (t < r.tmax) best_prim = prim;
            } else if (t < r.tmax || (t == r.tmax && best_prim != kInvalid && prim < best_prim)) {
                r.tmax = t;
                best_t = t;
                best_u = u;
                best_v = v;
                best_prim = prim;
                best_queue = __float_as_uint(b.w);
            }
        }
    };

    while (true) {
        // ---- refill idle lanes from the wave's chunk
        if (!exhausted) {
            const uint64_t idle = __ballot(!active);
            const uint32_t nidle = (uint32_t)__popcll(idle);
            if (nidle >= cfg.refill || nidle == 64u) {
                if (chunk_next >= chunk_end) {
                    // one atomic claims a chunk of rays for the whole wave
                    const uint32_t leader = (uint32_t)__ffsll((unsigned long long)idle) - 1u;
                    uint32_t base = 0;
                    if (lane_id() == leader) base = atomicAdd(work, chunk);
                    base = readlane_u32(base, leader);
                    base = static_end > n ? n : base + static_end;   // static_end >= n: nothing left
                    if (base >= n) {
                        exhausted = true;
                    } else {
                        chunk_next = base;
                        chunk_end = min(base + chunk, n);
                    }
                }
                if (chunk_next < chunk_end) {
                    const uint32_t avail = chunk_end - chunk_next;
                    const uint32_t lane_rank = mbcnt(idle);
                    if (!active && lane_rank < avail) {
                        V3 o, d;
                        float tmin, tmax;
                        uint32_t scene = 0;
                        item = src.load(chunk_next + lane_rank, o, d, tmin, tmax, scene);
                        if (MULTI) {
                            nodes_base = static_cast<const Node*>(NODE == kNodeBvh4Q ? scenes[scene].qnodes
                                                                                     : scenes[scene].nodes);
                            tris_base = scenes[scene].tris;
                        }
                        setup_ray(r, o, d, tmin, tmax);
                        best_prim = kInvalid;
                        best_queue = 0;
                        best_t = tmax;
                        best_u = best_v = 0.0f;
                        nodes = tris = 0;
                        ref = root;
                        sp = 0;
                        fin = false;
                        active = true;
                    }
                    chunk_next += min(nidle, avail);
                }
            }
        }
        if (__ballot(active) == 0) break;

        // ---- inner nodes: descend until every lane sits on a leaf or is done
        // (while-while); or one node step per outer iteration (if-if)
        while (true) {
            const bool step = active && !fin && !ref_is_leaf(ref);
            if (__ballot(step) == 0) break;
            if (step) {
                constexpr int W = NodeOf<NODE>::arity;
                float t0[W];
                uint32_t c[W];
                const Node* np = (kTopCap && (ref & kTopBit)) ? top + (ref & (kTopBit - 1u)) : nodes_base + ref;
                children(np, r, t0, c);
                if (COUNT) nodes++;
                if (!ANY || PT_ANY_SORTED) {
                    // nearest child first: sort the (distance, ref) pairs
                    if (W == 4) {
                        cswap(t0[0], c[0], t0[1], c[1]);
                        cswap(t0[W - 2], c[W - 2], t0[W - 1], c[W - 1]);
                        cswap(t0[0], c[0], t0[W - 2], c[W - 2]);
                        cswap(t0[1], c[1], t0[W - 1], c[W - 1]);
                        cswap(t0[1], c[1], t0[W - 2], c[W - 2]);
                    } else {
                        cswap(t0[0], c[0], t0[1], c[1]);
                    }
                }
                uint32_t next = kInvalid;
                float tnext = 0.0f;
#pragma unroll
                for (int i = W - 1; i >= 0; i--) {
                    if (t0[i] != __builtin_inff() && c[i] != kInvalid) {
                        if (next != kInvalid) push(next, tnext);   // the previous candidate is farther
                        next = c[i];
                        tnext = t0[i];
                    }
                }
                if (next == kInvalid) pop();
                else ref = next;
            }
            if (!PT_WHILE_WHILE) break;
        }

        // ---- leaves: test triangles two at a time, then continue from the stack
        while (true) {
            const bool step = active && !fin && ref_is_leaf(ref);
            if (__ballot(step) == 0) break;
            if (step) {
                const uint32_t first = leaf_first(ref);
                const uint32_t cnt = leaf_count(ref);
                if (COUNT) tris += cnt;
                for (uint32_t k = 0; k < cnt; k += 2) {
                    float4 a0, b0, c0, a1, b1, c1;
                    load_tri(tris_base + first + k, a0, b0, c0);
                    const bool two = k + 1 < cnt;
                    if (two) load_tri(tris_base + first + k + 1, a1, b1, c1);
                    test_tri(a0, b0, c0);
                    if (two) test_tri(a1, b1, c1);
                }
                if (ANY && best_prim != kInvalid) fin = true;
                else pop();
            }
            if (!PT_WHILE_WHILE) break;
        }

        // ---- finished rays write their result
        if (active && fin) {
            if (ANY) src.finish_any(item, best_prim != kInvalid, nodes, tris);
            else src.finish_closest(item, best_prim, best_t, best_u, best_v, best_queue, nodes, tris);
            acc_nodes += nodes;
            acc_tris += tris;
            active = false;
        }
    }
    if (COUNT) src.flush_counts(acc_nodes, acc_tris);
}

// Traversal of the compressed 8-wide nodes (Bvh8QNode), after the node-group
// traversal of Ylitie, Karras and Laine (HPG 2017).  A ray of octant oct (bit a
// set where rd_a < 0) takes a node's entered children in the visit order
// k ^ oct of their slots, which the builder's octant slots make close to
// nearest first: first the leaf children (one per outer iteration), then the
// inner children, whose pending ones stay on the stack as one group entry
// (first inner child, inner mask | pending visit mask << 8), so a node step
// pushes at most one entry and a pop reads the next child from the entry
// alone: no sort, no per-child push.  ANY: the first hit ends the ray;
// otherwise the closest hit, the (t, prim) minimum, whatever the order: every
// box is tested against the current closest t, inclusively, so a tie at that
// t is still reached.  Persistent lanes, chunked ray claims and the LDS stack
// with global spill as in trace_persistent; single-scene launches keep the
// first nodes (the top levels: breadth-first numbering) in LDS.
template <int BLOCK, int LDS, bool ANY, bool COUNT, bool MULTI, class Src>
__device__ __forceinline__ void trace_wide8(const Src& src, const SceneDev& sc, const SceneDev* scenes,
                                            const TraceConfig& cfg) {
    using Node = Bvh8QNode;
    __shared__ uint2 stack[LDS][BLOCK];
    constexpr uint32_t kTopCap =
        (MULTI || !PT_TOP_TABLE) ? 0u : (BLOCK >= 256 ? kTopBytes : (uint32_t)BLOCK * 32u) / (uint32_t)sizeof(Node);
    __shared__ Node top[kTopCap ? kTopCap : 1];
    const uint32_t tid = threadIdx.x;
    uint32_t top_n = 0;
    if (kTopCap && sc.qnodes8) {
        top_n = min(kTopCap, sc.node8_count);
        const uint4* s4 = static_cast<const uint4*>(sc.qnodes8);
        uint4* d4 = reinterpret_cast<uint4*>(top);
        const uint32_t words = top_n * (uint32_t)(sizeof(Node) / 16);
        for (uint32_t i = tid; i < words; i += BLOCK) d4[i] = s4[i];
        __syncthreads();
    }
    const uint32_t gtid = blockIdx.x * BLOCK + tid;
    const uint32_t stride = gridDim.x * BLOCK;
    const uint32_t cap = min(cfg.stack_cap, (uint32_t)LDS + cfg.spill_depth);
    const uint32_t n = src.count();
    unsigned int* work = src.work();
    const uint32_t waves = stride >> 6;
    const uint32_t chunk = min(max((n / (4u * waves) + 63u) & ~63u, 64u), max(cfg.chunk, 64u));

    bool active = false;      // lane holds a ray
    bool fin = false;         // its traversal has ended
    bool occluded = false;
    float best_t = 0.0f, best_u = 0.0f, best_v = 0.0f;      // closest hit so far (not ANY)
    uint32_t best_prim = kInvalid, best_queue = 0;
    const uint32_t wave_id = gtid >> 6;
    const uint32_t static_end = waves * chunk;
    uint32_t chunk_next = min(wave_id * chunk, n), chunk_end = min(wave_id * chunk + chunk, n);
    bool exhausted = chunk_next >= n;
    uint32_t item = 0, node = kInvalid, sp = 0, oct = 0;
    uint32_t lmask = 0, lbase = 0, lend0 = 0, lend1 = 0;   // leaf children of the last node, visit order
    const Node* nodes_base = static_cast<const Node*>(sc.qnodes8);
    const TriRecord* tris_base = sc.tris8;
    RayFrame r;
    uint32_t nodes = 0, tris = 0, acc_nodes = 0, acc_tris = 0;

    auto load_entry = [&](uint32_t i) -> uint2 {
        return i < (uint32_t)LDS ? stack[i][tid] : cfg.spill[(size_t)(i - LDS) * stride + gtid];
    };
    auto store_entry = [&](uint32_t i, uint2 e) {
        if (i < (uint32_t)LDS) stack[i][tid] = e;
        else cfg.spill[(size_t)(i - LDS) * stride + gtid] = e;
    };
    // the inner child at the lowest pending visit position of a group
    auto child_of = [&](uint32_t base, uint32_t imask, uint32_t pending) -> uint32_t {
        const uint32_t slot = (uint32_t)__builtin_ctz(pending) ^ oct;
        return base + (uint32_t)__builtin_popcount(imask & ((1u << slot) - 1u));
    };

    while (true) {
        // ---- refill idle lanes from the wave's chunk
        if (!exhausted) {
            const uint64_t idle = __ballot(!active);
            const uint32_t nidle = (uint32_t)__popcll(idle);
            if (nidle >= cfg.refill || nidle == 64u) {
                if (chunk_next >= chunk_end) {
                    const uint32_t leader = (uint32_t)__ffsll((unsigned long long)idle) - 1u;
                    uint32_t base = 0;
                    if (lane_id() == leader) base = atomicAdd(work, chunk);
                    base = readlane_u32(base, leader);
                    base = static_end > n ? n : base + static_end;
                    if (base >= n) {
                        exhausted = true;
                    } else {
                        chunk_next = base;
                        chunk_end = min(base + chunk, n);
                    }
                }
                if (chunk_next < chunk_end) {
                    const uint32_t avail = chunk_end - chunk_next;
                    const uint32_t lane_rank = mbcnt(idle);
                    if (!active && lane_rank < avail) {
                        V3 o, d;
                        float tmin, tmax;
                        uint32_t scene = 0;
                        item = src.load(chunk_next + lane_rank, o, d, tmin, tmax, scene);
                        if (MULTI) {
                            nodes_base = static_cast<const Node*>(scenes[scene].qnodes8);
                            tris_base = scenes[scene].tris8;
                        }
                        setup_ray(r, o, d, tmin, tmax);
                        oct = (r.rd.x >= 0.0f ? 0u : 1u) | (r.rd.y >= 0.0f ? 0u : 2u) | (r.rd.z >= 0.0f ? 0u : 4u);
                        occluded = false;
                        best_t = tmax;
                        best_u = best_v = 0.0f;
                        best_prim = kInvalid;
                        best_queue = 0;
                        nodes = tris = 0;
                        node = 0u;
                        lmask = 0u;
                        sp = 0;
                        fin = false;
                        active = true;
                    }
                    chunk_next += min(nidle, avail);
                }
            }
        }
        if (__ballot(active) == 0) break;

        // ---- one node step for lanes whose last node's leaves are done
        {
            const bool step = active && !fin && lmask == 0u && node != kInvalid;
            if (__ballot(step) != 0 && step) {
                const Node* np = node < top_n ? top + node : nodes_base + node;
                const Node8Hits h = enter8(np, r);
                if (COUNT) nodes++;
                // both masks into visit order at once: inner hits low byte, leaf hits high byte
                const uint32_t v = octant_order((h.hits & h.imask) | ((h.hits & ~h.imask) << 8), oct);
                const uint32_t inner = v & 0xFFu;
                lmask = v >> 8;
                lbase = h.tri_base;
                lend0 = h.end0;
                lend1 = h.end1;
                node = kInvalid;
                if (inner != 0u) {
                    node = child_of(h.node_base, h.imask, inner);
                    const uint32_t rest = inner & (inner - 1u);
                    if (rest != 0u) {
                        if (sp < cap) {
                            store_entry(sp, make_uint2(h.node_base, h.imask | (rest << 8)));
                            sp++;
                        } else {
                            atomicOr(cfg.overflow, 1u);
                        }
                    }
                }
            }
        }

        // ---- one leaf child: its triangles, two at a time
        {
            const bool step = active && !fin && lmask != 0u;
            if (__ballot(step) != 0 && step) {
                const uint32_t slot = (uint32_t)__builtin_ctz(lmask) ^ oct;
                lmask &= lmask - 1u;
                const uint32_t ends = slot < 4u ? lend0 : lend1;
                const uint32_t end = (ends >> (8u * (slot & 3u))) & 0xFFu;
                const uint32_t prev = slot == 0u ? 0u : slot == 4u ? (lend0 >> 24) : (ends >> (8u * ((slot - 1u) & 3u))) & 0xFFu;
                const uint32_t first = lbase + prev, cnt = end - prev;
                if (COUNT) tris += cnt;
                // the closest hit keeps the lower prim at equal t (Intersect.cu tie rule)
                auto closer = [&](const float4& a, const float4& b, float t, float u, float v) {
                    const uint32_t prim = __float_as_uint(a.w);
                    if (t < r.tmax || (t == r.tmax && best_prim != kInvalid && prim < best_prim)) {
                        r.tmax = t;
                        best_t = t;
                        best_u = u;
                        best_v = v;
                        best_prim = prim;
                        best_queue = __float_as_uint(b.w);
                    }
                };
                // ANY: an occluded ray stops testing the leaf (its count stays the leaf's)
                for (uint32_t k = 0; k < cnt && !(ANY && occluded); k += 2) {
                    float4 a0, b0, c0, a1, b1, c1;
                    load_tri(tris_base + first + k, a0, b0, c0);
                    const bool two = k + 1 < cnt;
                    if (two) load_tri(tris_base + first + k + 1, a1, b1, c1);
                    float t, u, v;
                    if (ANY) {
                        if (intersect_triangle(a0, b0, c0, r.o, r.d, r.tmin, t, u, v) && t < r.tmax) occluded = true;
                        if (two && intersect_triangle(a1, b1, c1, r.o, r.d, r.tmin, t, u, v) && t < r.tmax) occluded = true;
                    } else {
                        if (intersect_triangle(a0, b0, c0, r.o, r.d, r.tmin, t, u, v)) closer(a0, b0, t, u, v);
                        if (two && intersect_triangle(a1, b1, c1, r.o, r.d, r.tmin, t, u, v)) closer(a1, b1, t, u, v);
                    }
                }
                if (ANY && occluded) fin = true;
            }
        }

        // ---- lanes with nothing left at their node take the next child of the top group
        if (active && !fin && lmask == 0u && node == kInvalid) {
            if (sp == 0u) {
                fin = true;
            } else {
                const uint2 e = load_entry(sp - 1u);
                const uint32_t imask = e.y & 0xFFu, pending = e.y >> 8;
                node = child_of(e.x, imask, pending);
                const uint32_t rest = pending & (pending - 1u);
                if (rest == 0u) sp--;
                else store_entry(sp - 1u, make_uint2(e.x, imask | (rest << 8)));
            }
        }

        // ---- finished rays write their result
        if (active && fin) {
            if (ANY) src.finish_any(item, occluded, nodes, tris);
            else src.finish_closest(item, best_prim, best_t, best_u, best_v, best_queue, nodes, tris);
            acc_nodes += nodes;
            acc_tris += tris;
            active = false;
        }
    }
    if (COUNT) src.flush_counts(acc_nodes, acc_tris);
}

// ---------------------------------------------------------------- ray sources

struct WavefrontClosestSrc {
    PathState p;
    int cur;
    __device__ __forceinline__ uint32_t count() const { return p.ctr->ray[cur]; }
    __device__ __forceinline__ unsigned int* work() const { return &p.ctr->work_closest; }
    // rays are stored by queue position: consecutive lanes read consecutive records
    // ray_o.w holds the scene id; extension rays start at kMinT (PathTracerDefaults.hpp:13)
    __device__ __forceinline__ uint32_t load(uint32_t i, V3& o, V3& d, float& tmin, float& tmax, uint32_t& scene) const {
        const float4 a = (cur ? p.ray_o[1] : p.ray_o[0])[i];
        const float4 b = (cur ? p.ray_d[1] : p.ray_d[0])[i];
        o = xyz(a);
        scene = __float_as_uint(a.w);
        tmin = kMinT;
        d = xyz(b);
        tmax = b.w;
        return i;
    }
    // Hit and target queue of ray-queue entry i; the binning pass scatters
    // the entries into the material queues afterwards.
    __device__ __forceinline__ void finish_closest(uint32_t i, uint32_t prim, float t, float u, float v,
                                                   uint32_t queue, uint32_t, uint32_t) const {
        const bool hit = prim != kInvalid;
        if (hit) p.hit[i] = make_float4(t, u, v, __uint_as_float(prim));
        p.qid[i] = (uint8_t)(hit ? queue : (uint32_t)kQueueRegen);
    }
    __device__ void finish_any(uint32_t, bool, uint32_t, uint32_t) const {}
    __device__ __forceinline__ void flush_counts(uint32_t nodes, uint32_t tris) const {
        add_counts(&p.ctr->closest_nodes, &p.ctr->closest_tris, nodes, tris);
    }
};

// Adds the contribution L that a path made at a later vertex to every guide
// record of the path: record k receives L / (path throughput after vertex k),
// the radiance that arrived at vertex k along its sampled direction.
__device__ __forceinline__ void guide_credit(GuideRecord* recs, uint32_t head, V3 L) {
    uint32_t r = head;
    for (uint32_t n = 0; r != kInvalid && n <= kMaxDepthLimit; n++) {
        GuideRecord& g = recs[r];
        const float4 t = *reinterpret_cast<const float4*>(g.t_after);
        float4 li = *reinterpret_cast<const float4*>(g.li);
        li.x += t.x > 0.0f ? L.x / t.x : 0.0f;
        li.y += t.y > 0.0f ? L.y / t.y : 0.0f;
        li.z += t.z > 0.0f ? L.z / t.z : 0.0f;
        *reinterpret_cast<float4*>(g.li) = li;
        r = g.prev;
    }
}

struct WavefrontShadowSrc {
    PathState p;
    int cur;   // parity of the iteration whose shade pushed the rays
    __device__ __forceinline__ uint32_t count() const { return p.ctr->shadow[cur]; }
    __device__ __forceinline__ unsigned int* work() const { return &p.ctr->work_shadow[cur]; }
    __device__ __forceinline__ uint32_t load(uint32_t i, V3& o, V3& d, float& tmin, float& tmax, uint32_t& scene) const {
        const float4 a = p.sh_o[i], b = p.sh_d[i];
        o = xyz(a);
        tmax = a.w;
        d = xyz(b);
        scene = __float_as_uint(b.w);
        tmin = kMinT;
        return i;
    }
    __device__ void finish_closest(uint32_t, uint32_t, float, float, float, uint32_t, uint32_t, uint32_t) const {}
    __device__ __forceinline__ void finish_any(uint32_t i, bool occluded, uint32_t, uint32_t) const {
        if (!occluded) {
            const float4 L4 = p.sh_L[i];
            V3 L = xyz(L4);
            // ClampRadiance(radiance, 20), Intersect.cu:125-134
            const float l2 = dot(L, L);
            const float max_len = 20.0f;
            if (l2 > max_len * max_len) L = L * (max_len * (1.0f / sqrtf(l2)));
            float* acc = reinterpret_cast<float*>(p.accum + __float_as_uint(L4.w));
            atomicAdd(acc + 0, L.x);
            atomicAdd(acc + 1, L.y);
            atomicAdd(acc + 2, L.z);
            if (p.guide.sh_head) guide_credit(p.guide.recs, p.guide.sh_head[i], L);
        }
    }
    __device__ __forceinline__ void flush_counts(uint32_t nodes, uint32_t tris) const {
        add_counts(&p.ctr->shadow_nodes, &p.ctr->shadow_tris, nodes, tris);
    }
};

struct BatchSrc {
    const float4* rays;
    float4* hits;
    uint32_t* occluded;
    uint32_t* nodes_out;
    uint32_t* tris_out;
    unsigned int* work_ctr;
    uint32_t n;
    uint8_t* qid;                   // optional: target queue of each ray
    __device__ __forceinline__ uint32_t count() const { return n; }
    __device__ __forceinline__ unsigned int* work() const { return work_ctr; }
    __device__ __forceinline__ uint32_t load(uint32_t i, V3& o, V3& d, float& tmin, float& tmax, uint32_t& scene) const {
        scene = 0;
        const float4 a = rays[2 * (size_t)i], b = rays[2 * (size_t)i + 1];
        o = xyz(a);
        tmin = a.w;
        d = xyz(b);
        tmax = b.w;
        return i;
    }
    __device__ __forceinline__ void finish_closest(uint32_t i, uint32_t prim, float t, float u, float v,
                                                   uint32_t queue, uint32_t nodes, uint32_t tris) const {
        // pt_hit layout: prim, t, u, v
        if (hits) {
            if (prim == kInvalid) hits[i] = make_float4(__uint_as_float(kInvalid), 0.0f, 0.0f, 0.0f);
            else hits[i] = make_float4(__uint_as_float(prim), t, u, v);
        }
        if (qid) qid[i] = (uint8_t)(prim != kInvalid ? queue : (uint32_t)kQueueRegen);
        if (nodes_out) nodes_out[i] = nodes;
        if (tris_out) tris_out[i] = tris;
    }
    __device__ void flush_counts(uint32_t, uint32_t) const {}
    __device__ __forceinline__ void finish_any(uint32_t i, bool occ, uint32_t nodes, uint32_t tris) const {
        occluded[i] = occ ? 1u : 0u;
        if (nodes_out) nodes_out[i] = nodes;
        if (tris_out) tris_out[i] = tris;
    }
};

template <int BLOCK, int LDS, int W, bool ANY, bool COUNT, bool MULTI, class Src>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(COUNT ? 1 : PT_TRACE_WAVES)))
void k_trace(Src src, SceneDev sc, const SceneDev* scenes, TraceConfig cfg) {
    if constexpr (W == kNodeBvh8Q) {
        trace_wide8<BLOCK, LDS, ANY, COUNT, MULTI, Src>(src, sc, scenes, cfg);
    } else {
        trace_persistent<BLOCK, LDS, W, ANY, COUNT, MULTI, Src>(src, sc, scenes, cfg);
    }
}

// ================================================================ shading math

struct Frame3 {
    V3 t, b, n;
    __device__ __forceinline__ V3 to_local(V3 d) const { return v3(dot(d, t), dot(d, b), dot(d, n)); }
    __device__ __forceinline__ V3 from_local(V3 d) const { return d.x * t + d.y * b + d.z * n; }
};

// Math::SafeNormalize (Math.hpp:108-118) with 1/sqrt in place of rsqrt.
__device__ __forceinline__ V3 safe_normalize(V3 v, V3 fallback) {
    const float l2 = dot(v, v);
    if (l2 > 1e-16f) return v * (1.0f / sqrtf(l2));
    return fallback;
}

// Math::BuildFrame(normal) (Math.hpp:120-134)
__device__ __forceinline__ Frame3 build_frame(V3 n) {
    V3 t;
    if (fabsf(n.x) > fabsf(n.z)) t = normalize(v3(n.y, -n.x, 0.0f));
    else t = normalize(v3(0.0f, n.z, -n.y));
    return Frame3{t, cross(n, t), n};
}

// Math::BuildFrame(normal, tangent, sign) (Math.hpp:136-148)
__device__ __forceinline__ Frame3 build_frame(V3 n, V3 t, float sign) {
    n = safe_normalize(n, v3(0.0f, 0.0f, 1.0f));
    const Frame3 fb = build_frame(n);
    t = t - n * dot(n, t);
    t = safe_normalize(t, fb.t);
    return Frame3{t, cross(n, t) * sign, n};
}

__device__ __forceinline__ V3 fix_normal(V3 gn, V3 d) { return dot(gn, d) > 0.0f ? -gn : gn; }

__device__ __forceinline__ float luminance(V3 c) { return 0.2126f * c.x + 0.7152f * c.y + 0.0722f * c.z; }

// sin and cos of a sampling angle x = 2 pi u, 0 <= x < 2 pi (the reference
// calls the device libm, MaterialEval.cu:228,591).  Quadrant reduction by
// pi/2 split into three parts (Cody & Waite 1980; the first two parts have 17
// significant bits, so q * part is exact for q <= 4) and the Taylor
// polynomials of degree 9 and 10 on |r| <= pi/4: at most 1.6 ulp from the
// exact values, and with contraction off every operation rounds as the CPU
// oracle's, so a sample's directions are identical on both.
__device__ __forceinline__ void sincos_angle(float x, float* s, float* c) {
    constexpr float kTwoOverPi = 0x1.45f306p-1f;
    constexpr float kP1 = 0x1.921ep+0f, kP2 = 0x1.b544p-16f, kP3 = 0x1.0b4612p-34f;
    const float q = rintf(x * kTwoOverPi);
    float r = x - q * kP1;
    r = r - q * kP2;
    r = r - q * kP3;
    const float r2 = r * r;
    const float sp = r + r * (r2 * (-0x1.555556p-3f + r2 * (0x1.111112p-7f + r2 * (-0x1.a01a02p-13f + r2 * 0x1.71de3ap-19f))));
    const float cp = (1.0f - 0.5f * r2) +
                     (r2 * r2) * (0x1.555556p-5f + r2 * (-0x1.6c16c2p-10f + r2 * (0x1.a01a02p-16f + r2 * -0x1.27e4fcp-22f)));
    const int k = (int)q & 3;
    const float a = (k & 1) ? cp : sp, b = (k & 1) ? sp : cp;
    *s = (k & 2) ? -a : a;
    *c = ((k + 1) & 2) ? -b : b;
}

__device__ __forceinline__ V3 cosine_hemisphere(float u0, float u1) {
    const float phi = 2.0f * kPi * u0;
    const float cos_t = sqrtf(u1);
    const float sin_t = sqrtf(1.0f - u1);
    float s, c;
    sincos_angle(phi, &s, &c);
    return v3(sin_t * c, sin_t * s, cos_t);
}

// Math.hpp:217-262 trig helpers on a local direction
__device__ __forceinline__ float cos2_theta(V3 w) { return w.z * w.z; }
__device__ __forceinline__ float sin2_theta(V3 w) { return fmaxf(0.0f, 1.0f - cos2_theta(w)); }
__device__ __forceinline__ float sin_theta(V3 w) { return sqrtf(sin2_theta(w)); }
__device__ __forceinline__ float tan2_theta(V3 w) { return sin2_theta(w) / cos2_theta(w); }
__device__ __forceinline__ float cos_phi(V3 w) {
    const float st = sin_theta(w);
    return st == 0.0f ? 1.0f : clampf(w.x / st, -1.0f, 1.0f);
}
__device__ __forceinline__ float sin_phi(V3 w) {
    const float st = sin_theta(w);
    return st == 0.0f ? 0.0f : clampf(w.y / st, -1.0f, 1.0f);
}

// GGX terms, MaterialEval.cu:587-650
__device__ __forceinline__ float ggx_lambda(V3 w, float alpha) {
    const float t2 = tan2_theta(w);
    if (t2 == __builtin_inff()) return 0.0f;
    const float a2 = sqr(cos_phi(w) * alpha) + sqr(sin_phi(w) * alpha);
    return (sqrtf(1.0f + a2 * t2) - 1.0f) * 0.5f;
}
__device__ __forceinline__ float ggx_g1(V3 w, float alpha) { return 1.0f / (1.0f + ggx_lambda(w, alpha)); }
__device__ __forceinline__ float ggx_g(V3 wo, V3 wi, float alpha) {
    return 1.0f / (1.0f + ggx_lambda(wo, alpha) + ggx_lambda(wi, alpha));
}
__device__ __forceinline__ float ggx_d(V3 wm, float alpha) {
    const float t2 = tan2_theta(wm);
    if (t2 == __builtin_inff()) return 0.0f;
    const float cos4 = sqr(cos2_theta(wm));
    if (cos4 < 1e-16f) return 0.0f;
    const float e = t2 * (sqr(cos_phi(wm) / alpha) + sqr(sin_phi(wm) / alpha));
    return 1.0f / (kPi * sqr(alpha) * cos4 * sqr(1.0f + e));
}
__device__ __forceinline__ float ggx_dv(V3 w, V3 wm, float alpha) {
    return ggx_g1(w, alpha) / fabsf(w.z) * ggx_d(wm, alpha) * fabsf(dot(w, wm));
}
// FresnelSchlick, MaterialEval.cu:647-650, with (1 - c)^5 as three products
__device__ __forceinline__ V3 fresnel_schlick(V3 f0, float cos_h) {
    const float x = 1.0f - cos_h;
    const float x2 = x * x;
    const float k = (x2 * x2) * x;
    return f0 + (v3(1.0f) - f0) * k;
}
// Visible-normal sampling (Heitz 2018), TrowbridgeSampleWm, MaterialEval.cu:594-610
__device__ __forceinline__ V3 ggx_sample_wm(V3 w, float alpha, float u0, float u1) {
    V3 wh = normalize(v3(alpha * w.x, alpha * w.y, w.z));
    if (wh.z < 0.0f) wh = -wh;
    const V3 t1 = (wh.z < 0.999f) ? normalize(cross(wh, v3(0.0f, 0.0f, 1.0f))) : v3(1.0f, 0.0f, 0.0f);
    const V3 t2 = cross(wh, t1);
    const float rr = sqrtf(u0);
    const float theta = 2.0f * kPi * u1;
    float s, c;
    sincos_angle(theta, &s, &c);
    float px = rr * c, py = rr * s;
    const float h = sqrtf(1.0f - px * px);
    const float lt = (1.0f + wh.z) * 0.5f;
    py = h + lt * (py - h);
    const float pz = sqrtf(fmaxf(0.0f, 1.0f - (px * px + py * py)));
    const V3 nh = px * t1 + py * t2 + pz * wh;
    return normalize(v3(alpha * nh.x, alpha * nh.y, fmaxf(1e-6f, nh.z)));
}

// ---------------------------------------------------------------- per-sample state

struct Draws {
    uint64_t state, inc;
    __device__ __forceinline__ float next() {
        const uint32_t x = pcg_output(state);
        state = state * kPcgMult + inc;
        return u01(x);
    }
};

__device__ __forceinline__ Draws draws_at(const FrameParams* fp, uint64_t seed, uint64_t sample, uint32_t bounce) {
    const PcgStream s = pcg_seed(seed, sample);
    const PcgJump j = fp->bounce_jump[bounce];
    return Draws{j.mult * s.state + j.plus * s.inc, s.inc};
}

struct Hit {
    V3 p, gn, n;
    float4 tan;
};

__device__ __forceinline__ float4 lerp4(float b0, float4 a, float b1, float4 b, float b2, float4 c) {
    return make_float4(b0 * a.x + b1 * b.x + b2 * c.x, b0 * a.y + b1 * b.y + b2 * c.y,
                       b0 * a.z + b1 * b.z + b2 * c.z, b0 * a.w + b1 * b.w + b2 * c.w);
}

// LoadResolvedHit, MaterialEval.cu:54-87 (no UVs: materials are constant)
__device__ __forceinline__ Hit resolve_hit(const ShadeRecord* rec, V3 o, V3 d, float t, float u, float v,
                                           uint32_t& material) {
    const float4* s = reinterpret_cast<const float4*>(rec);
    const float4 n0 = s[0], n1 = s[1], n2 = s[2], t0 = s[3], t1 = s[4], t2 = s[5], g = s[6];
    const float b0 = 1.0f - u - v, b1 = u, b2 = v;
    Hit h;
    h.p = o + d * t;
    h.gn = fix_normal(xyz(g), d);
    h.n = b0 * xyz(n0) + b1 * xyz(n1) + b2 * xyz(n2);
    h.tan = lerp4(b0, t0, b1, t1, b2, t2);
    material = __float_as_uint(g.w);
    return h;
}

// GetShadingFrame, MaterialEval.cu:96-115, with the normal-map texel as a material constant.
__device__ __forceinline__ Frame3 shading_frame(const Hit& h, V3 nts) {
    const V3 bt = v3(h.tan.x, h.tan.y, h.tan.z);
    const float sign = h.tan.w < 0.0f ? -1.0f : 1.0f;
    const Frame3 base = build_frame(h.n, bt, sign);
    const V3 sn = nts.x * base.t + nts.y * base.b + nts.z * base.n;
    return build_frame(sn, bt, sign);
}

struct Continuation {
    bool terminated;
    uint32_t depth;
    V3 thr;
};

// ComputePathContinuation, MaterialEval.cu:129-146.  rr_scale multiplies the
// throughput that sets the survival probability; it is exactly 1 for plain
// BSDF sampling, which keeps the reference's roulette bit for bit.
__device__ __forceinline__ Continuation continue_path(uint32_t depth, V3 thr, uint32_t limit, uint32_t rr_start,
                                                      V3 weight, float rr, float rr_scale = 1.0f) {
    const uint32_t next = depth + 1u;
    if (next >= limit) return Continuation{true, next, v3(0.0f)};
    const V3 nt = thr * weight;
    if (next < rr_start) return Continuation{false, next, nt};
    const float q = clampf(maxc(nt) * rr_scale, 0.0f, 1.0f);
    if (rr > q) return Continuation{true, next, v3(0.0f)};
    return Continuation{false, next, nt / q};
}

struct LightSample {
    V3 p, gn, wi;
    float d2;
    V3 emission;
    float pdf;
};

// AliasTableView::Sample, AliasTableView.hpp:32-60
__device__ __forceinline__ uint32_t alias_pick(const AliasBin* bins, uint32_t size, float u, float& pmf) {
    const float scaled = u * (float)size;
    const uint32_t bi = min((uint32_t)scaled, size - 1u);
    const float up = fminf(scaled - (float)bi, kOneMinusEpsilon);
    const AliasBin b = bins[bi];
    if (up < b.q) {
        pmf = b.p;
        return bi;
    }
    pmf = bins[b.alias].p;
    return b.alias;
}

// SampleLightTriangle, MaterialEval.cu:171-192
__device__ __forceinline__ LightSample sample_light(const SceneDev& sc, float us, float x0, float x1, V3 p) {
    float pmf;
    const uint32_t li = alias_pick(sc.bins, sc.light_count, us, pmf);
    const float4* L = reinterpret_cast<const float4*>(sc.lights + li);
    const float4 a = L[0], e1 = L[1], e2 = L[2], g = L[3], em = L[4];
    const float s0 = sqrtf(x0);
    const float b1 = s0 * (1.0f - x1);
    const float b2 = s0 * x1;
    LightSample ls;
    ls.p = xyz(a) + b1 * xyz(e1) + b2 * xyz(e2);
    const V3 to = ls.p - p;
    ls.d2 = dot(to, to);
    ls.wi = normalize(to);
    ls.gn = fix_normal(xyz(g), ls.wi);
    ls.emission = xyz(em);
    ls.pdf = pmf / a.w;
    return ls;
}

struct ShadowOut {
    bool emit;
    V3 o, d, L;
    float tmax;
};

__device__ __forceinline__ void make_shadow(ShadowOut& so, V3 p, V3 gn, V3 lp, V3 L) {
    const V3 origin = p + gn * 1e-4f;
    const V3 sv = lp - origin;
    const float dist = length(sv);
    so.emit = true;
    so.o = origin;
    so.d = sv / dist;
    so.tmax = dist - 1e-4f;
    so.L = L;
}

// ================================================================ kernels

__global__ void k_frame_begin(PathState p) {
    if (threadIdx.x == 0) {
        Counters* c = p.ctr;
        const FrameParams* fp = p.fp;
        const uint64_t total = fp->total_samples;
        const uint32_t gen = (uint32_t)(total < (uint64_t)p.capacity ? total : (uint64_t)p.capacity);
        c->ray[0] = gen;
        c->ray[1] = 0;
        for (int i = 0; i < kQueueSlots; i++) c->queue[i] = 0;
        c->shadow[0] = c->shadow[1] = 0;
        c->work_closest = 0;
        c->work_shadow[0] = c->work_shadow[1] = 0;
        c->active = gen;
        c->iterations = 0;
        c->invalid_pixels = 0;
        c->stack_overflow = 0;
        c->records = 0;
        c->launched = gen;
        c->ext_rays = 0;
        c->shadow_rays = 0;
        c->closest_nodes = c->closest_tris = 0;
        c->shadow_nodes = c->shadow_tris = 0;
        c->dg_entries = 0;
    }
}

__global__ void k_fill_accum(float4* accum, uint32_t begin, uint32_t end) {
    for (uint32_t i = begin + blockIdx.x * blockDim.x + threadIdx.x; i < end; i += gridDim.x * blockDim.x)
        accum[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

// Ring slot of the view owning global sample g: the last live view whose
// sample_begin <= g (virtual indices in [lo, hi), sample_begin increasing).
__device__ __forceinline__ uint32_t view_of_sample(const ViewRecord* views, uint32_t lo, uint32_t hi, uint64_t g) {
    while (hi - lo > 1) {
        const uint32_t mid = lo + ((hi - lo) >> 1);
        if (views[mid & (kViewRing - 1u)].sample_begin <= g) lo = mid;
        else hi = mid;
    }
    return lo & (kViewRing - 1u);
}

// InitializePathsKernel / RegeneratePathsKernel, PathGeneration.cu:56-102.
// New paths go to the free tail of the next ray queue: positions
// [ray[nxt], ray[nxt] + n) with n = min(free positions, samples left).
// Global sample g belongs to the view record in ring slot v = view_of_sample(g);
// its local index g - sample_begin + sample_offset addresses the view's pixels
// and sample streams.
__global__ void k_generate(PathState p, int init, int nxt) {
    const FrameParams* fp = p.fp;
    Counters* c = p.ctr;
    uint32_t n, base;
    uint64_t launched;
    if (init) {
        n = c->ray[0];
        base = 0;
        launched = 0;
    } else {
        launched = c->launched;
        const uint64_t remaining = fp->total_samples - launched;
        base = c->ray[nxt];
        const uint32_t free_slots = p.capacity - base;
        n = (uint64_t)free_slots < remaining ? free_slots : (uint32_t)remaining;
    }
    const uint32_t view_lo = fp->view_lo, view_hi = fp->view_hi;
    float4* ro = nxt ? p.ray_o[1] : p.ray_o[0];
    float4* rd = nxt ? p.ray_d[1] : p.ray_d[0];
    float4* th = nxt ? p.thr[1] : p.thr[0];
    uint4* me = nxt ? p.meta[1] : p.meta[0];
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        const uint64_t g = launched + k;
        const uint32_t v = view_of_sample(p.views, view_lo, view_hi, g);
        const ViewRecord* vr = p.views + v;
        const uint64_t sample = g - vr->sample_begin + vr->sample_offset;
        const uint32_t W = vr->width, H = vr->height, spp = vr->spp, axis = vr->spp_axis;
        const PcgStream s = pcg_seed(vr->seed, sample);
        Draws dr{s.state, s.inc};
        const float jx = dr.next();
        const float jy = dr.next();
        // ComputePathSample, PathGeneration.cu:15-33
        const uint64_t pixel64 = sample / spp;
        const uint32_t cell = (uint32_t)(sample - pixel64 * spp);
        const uint32_t pixel = (uint32_t)pixel64;
        const uint32_t y = pixel / W, x = pixel - y * W;
        const uint32_t cy = cell / axis, cx = cell - cy * axis;
        const float u = ((float)x + ((float)cx + jx) / (float)axis) / (float)W;
        const float vv = ((float)y + ((float)cy + jy) / (float)axis) / (float)H;
        // GenerateCameraRay, PathGeneration.cu:35-43
        const V3 co = v3(vr->cam.origin[0], vr->cam.origin[1], vr->cam.origin[2]);
        const V3 llc = v3(vr->cam.llc[0], vr->cam.llc[1], vr->cam.llc[2]);
        const V3 hor = v3(vr->cam.horizontal[0], vr->cam.horizontal[1], vr->cam.horizontal[2]);
        const V3 ver = v3(vr->cam.vertical[0], vr->cam.vertical[1], vr->cam.vertical[2]);
        const V3 d = normalize(llc + u * hor + vv * ver - co);
        const uint32_t pos = base + k;
        ro[pos] = f4(co, __uint_as_float(vr->scene));
        rd[pos] = f4(d, kMaxT);
        th[pos] = make_float4(1.0f, 1.0f, 1.0f, 1.0f);
        me[pos] = make_uint4(pixel, kDeltaBit | (v << 8), (uint32_t)sample, (uint32_t)(sample >> 32));
        if (p.guide.head[0]) {
            (nxt ? p.guide.head[1] : p.guide.head[0])[pos] = kInvalid;
            (nxt ? p.guide.rr_scale[1] : p.guide.rr_scale[0])[pos] = 1.0f;
        }
    }
}

// PrepareIterationKernel (Iteration.cu:7-23) plus the host-side bookkeeping of
// Renderer.cpp:573-619, done on the device so the loop needs no host sync.
__global__ void k_prepare(PathState p, int cur) {
    if (threadIdx.x == 0) {
        Counters* c = p.ctr;
        const int nxt = cur ^ 1;
        const uint64_t remaining = p.fp->total_samples - c->launched;
        const uint32_t free_slots = p.capacity - c->ray[nxt];
        const uint32_t nr = (uint64_t)free_slots < remaining ? free_slots : (uint32_t)remaining;
        const uint32_t traced = c->ray[cur];
        c->ext_rays += traced;
        c->shadow_rays += c->shadow[cur];
        if (traced) c->iterations++;
        c->dg_entries += (unsigned long long)c->queue[0] + c->queue[1];
        c->ray[nxt] += nr;
        c->launched += nr;
        c->ray[cur] = 0;
        for (int i = 0; i < kQueueSlots; i++) c->queue[i] = 0;
        // this iteration's shadow rays may still be in flight on another
        // stream: clear the pair the next shade pushes into
        c->shadow[nxt] = 0;
        c->work_closest = 0;
        c->work_shadow[nxt] = 0;
        c->active = c->ray[nxt];
    }
}

// Scatters entries [0, n) into the five queues by their queue id.  Each block
// owns one contiguous range: a counting pass, one atomicAdd per queue per
// block, then a scatter pass in which every wave writes at its own running
// offsets.  Five global atomics per block in place of one per finishing ray.
// Both passes take kBinUnroll rows of the block per step, their queue ids
// loaded together (independent loads in flight), then handled in index order.
constexpr int kBinBlock = 256;
constexpr int kBinWaves = kBinBlock / 64;
constexpr int kBinUnroll = 4;

template <class SlotOf>
__device__ __forceinline__ void bin_entries(const uint8_t* qid, uint32_t n, const SlotOf& slot_of,
                                            uint32_t* const* queues, unsigned int* counters) {
    __shared__ uint32_t s_cnt[kQueueSlots][kBinWaves];
    const uint32_t per = (((n + gridDim.x - 1) / gridDim.x) + 63u) & ~63u;
    const uint32_t b0 = blockIdx.x * per;
    const uint32_t b1 = min(b0 + per, n);
    const uint32_t wave = threadIdx.x >> 6, lane = lane_id();
    uint32_t cnt[kQueueSlots] = {0, 0, 0, 0, 0};
    for (uint32_t base = b0; base < b1; base += kBinUnroll * kBinBlock) {
        int qs[kBinUnroll];
#pragma unroll
        for (int u = 0; u < kBinUnroll; u++) {
            const uint32_t i = base + (uint32_t)u * kBinBlock + threadIdx.x;
            qs[u] = i < b1 ? (int)qid[i] : -1;
        }
#pragma unroll
        for (int u = 0; u < kBinUnroll; u++)
#pragma unroll
            for (int k = 0; k < kQueueSlots; k++) cnt[k] += (uint32_t)__popcll(__ballot(qs[u] == k));
    }
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < kQueueSlots; k++) s_cnt[k][wave] = cnt[k];
    }
    __syncthreads();
    if (threadIdx.x < (uint32_t)kQueueSlots) {
        const int k = (int)threadIdx.x;
        uint32_t total = 0;
        for (int w = 0; w < kBinWaves; w++) total += s_cnt[k][w];
        uint32_t run = total ? atomicAdd(counters + k, total) : 0u;
        for (int w = 0; w < kBinWaves; w++) {
            const uint32_t c = s_cnt[k][w];
            s_cnt[k][w] = run;
            run += c;
        }
    }
    __syncthreads();
    uint32_t off[kQueueSlots];
#pragma unroll
    for (int k = 0; k < kQueueSlots; k++) off[k] = s_cnt[k][wave];
    for (uint32_t base = b0; base < b1; base += kBinUnroll * kBinBlock) {
        int qs[kBinUnroll];
#pragma unroll
        for (int u = 0; u < kBinUnroll; u++) {
            const uint32_t i = base + (uint32_t)u * kBinBlock + threadIdx.x;
            qs[u] = i < b1 ? (int)qid[i] : -1;
        }
#pragma unroll
        for (int u = 0; u < kBinUnroll; u++) {
            const uint32_t i = base + (uint32_t)u * kBinBlock + threadIdx.x;
            const int q = qs[u];
            const uint32_t slot = q >= 0 ? slot_of(i) : 0u;
#pragma unroll
            for (int k = 0; k < kQueueSlots; k++) {
                const uint64_t m = __ballot(q == k);
                if (q == k && queues[k]) queues[k][off[k] + mbcnt(m)] = slot;
                off[k] += (uint32_t)__popcll(m);
            }
        }
    }
}

struct IdentitySlot {
    __device__ __forceinline__ uint32_t operator()(uint32_t i) const { return i; }
};

__global__ __launch_bounds__(kBinBlock) void k_bin(PathState p, int cur) {
    const uint32_t n = p.ctr->ray[cur];
    uint32_t* queues[kQueueSlots] = {p.q[0], p.q[1], p.q[2], p.q[3], nullptr};
    bin_entries(p.qid, n, IdentitySlot{}, queues, p.ctr->queue);
}

struct QueueSet {
    uint32_t* q[kQueueSlots];
};

__global__ __launch_bounds__(kBinBlock) void k_bin_batch(const uint8_t* qid, uint32_t n, QueueSet qs,
                                                          unsigned int* counters) {
    bin_entries(qid, n, IdentitySlot{}, qs.q, counters);
}

// Block-wide reservation for two output queues: one atomicAdd per queue per
// block.  Every thread of the block must call it.
constexpr uint32_t kMaxBlockWaves = 16;   // largest block of the kernels that reserve queue space

__device__ __forceinline__ void block_reserve2(bool pred0, bool pred1, unsigned int* c0, unsigned int* c1,
                                               uint32_t& pos0, uint32_t& pos1) {
    __shared__ uint32_t s_off[2][kMaxBlockWaves];
    const uint32_t wave = threadIdx.x >> 6, lane = lane_id();
    const uint32_t waves = blockDim.x >> 6;
    const uint64_t m0 = __ballot(pred0), m1 = __ballot(pred1);
    if (lane == 0) {
        s_off[0][wave] = (uint32_t)__popcll(m0);
        s_off[1][wave] = (uint32_t)__popcll(m1);
    }
    __syncthreads();
    if (threadIdx.x < 2u) {
        const int k = (int)threadIdx.x;
        unsigned int* ctr = k == 0 ? c0 : c1;
        uint32_t total = 0;
        for (uint32_t w = 0; w < waves; w++) total += s_off[k][w];
        uint32_t run = total ? atomicAdd(ctr, total) : 0u;
        for (uint32_t w = 0; w < waves; w++) {
            const uint32_t c = s_off[k][w];
            s_off[k][w] = run;
            run += c;
        }
    }
    __syncthreads();
    pos0 = s_off[0][wave] + mbcnt(m0);
    pos1 = s_off[1][wave] + mbcnt(m1);
    __syncthreads();
}

// block_reserve2 with a third counter (guide training records).
__device__ __forceinline__ void block_reserve3(bool pred0, bool pred1, bool pred2, unsigned int* c0, unsigned int* c1,
                                               unsigned int* c2, uint32_t& pos0, uint32_t& pos1, uint32_t& pos2) {
    __shared__ uint32_t s_off3[3][kMaxBlockWaves];
    const uint32_t wave = threadIdx.x >> 6, lane = lane_id();
    const uint32_t waves = blockDim.x >> 6;
    const uint64_t m0 = __ballot(pred0), m1 = __ballot(pred1), m2 = __ballot(pred2);
    if (lane == 0) {
        s_off3[0][wave] = (uint32_t)__popcll(m0);
        s_off3[1][wave] = (uint32_t)__popcll(m1);
        s_off3[2][wave] = (uint32_t)__popcll(m2);
    }
    __syncthreads();
    if (threadIdx.x < 3u) {
        const int k = (int)threadIdx.x;
        unsigned int* ctr = k == 0 ? c0 : k == 1 ? c1 : c2;
        uint32_t total = 0;
        for (uint32_t w = 0; w < waves; w++) total += s_off3[k][w];
        uint32_t run = total ? atomicAdd(ctr, total) : 0u;
        for (uint32_t w = 0; w < waves; w++) {
            const uint32_t c = s_off3[k][w];
            s_off3[k][w] = run;
            run += c;
        }
    }
    __syncthreads();
    pos0 = s_off3[0][wave] + mbcnt(m0);
    pos1 = s_off3[1][wave] + mbcnt(m1);
    pos2 = s_off3[2][wave] + mbcnt(m2);
    __syncthreads();
}

// The four material stages of one bounce in one launch: element e of the
// virtual range [diffuse | ggx | mirror | emissive] is queue entry e - offset.
// Each wave walks 64 consecutive entries, so a wave shades one material except
// at the three queue boundaries.  MULTI: paths of several views (per-lane view
// and scene records); otherwise every path belongs to view 0 of scene `sc1`,
// whose records are wave-uniform loads.  GUIDED: diffuse and GGX directions by
// one-sample MIS of the guide mixture (probability cfg.guide_fraction) and the
// BSDF, each continuing diffuse/GGX vertex writes a training record, and
// radiance reaching the pixel is credited to the path's records.
template <bool MULTI, bool GUIDED>
__global__ __launch_bounds__(GUIDED ? 64 * kInferWaves : 256) void k_shade(PathState p, SceneDev sc1, int cur) {
    constexpr uint32_t kBlock = GUIDED ? 64u * kInferWaves : 256u;
    const FrameParams* fp = p.fp;
    Counters* c = p.ctr;
    const int nxt = cur ^ 1;
    const uint32_t c0 = c->queue[0], c1 = c->queue[1], c2 = c->queue[2], c3 = c->queue[3];
    const uint32_t e1 = c0, e2 = c0 + c1, e3 = e2 + c2, total = e3 + c3;
    const uint32_t rec_mask = GUIDED ? p.guide.cfg.rec_mask : 0u;
    const PcgJump jump7 = pcg_jump(7u);
    const float gfrac = GUIDED ? p.guide.cfg.guide_fraction : 0.0f;
    // GUIDED: the guide network in LDS; each wave infers the mixtures of its 64
    // entries into its own tile, which the lanes then read while shading
    InferLds<kInferWaves, 3>* gl = nullptr;
    if constexpr (GUIDED) {
        __shared__ InferLds<kInferWaves, 3> glds;
        gl = &glds;
        if (gfrac > 0.0f) stage_weights(gl->w, gl->b, p.guide.wf, p.guide.bf, InferLds<kInferWaves, 3>::kWeights);
    }

    for (uint32_t bbase = blockIdx.x * kBlock; bbase < total; bbase += gridDim.x * kBlock) {
        const uint32_t e = bbase + threadIdx.x;
        const bool valid = e < total;
        uint32_t* mixw = nullptr;          // this wave's mixture words: word k of lane s at mixw[64 k + s]
        if constexpr (GUIDED) {
            if (gfrac > 0.0f) {
                const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
                bf16x8* xt = gl->x[wave];
                mixw = reinterpret_cast<uint32_t*>(xt);
                const bool need = valid && e < e2;
                if (__ballot(need) != 0) {
                    float f[kGuideDim];
#pragma unroll
                    for (int k = 0; k < kGuideDim; k++) f[k] = 0.0f;
                    if (need) InferPathSrc{p, sc1, cur}.features(e, f);
                    store_input_frags(xt, lane, f);
                    wave_lds_sync();
                    f32x4 o[4][4];
                    tile_forward<3>(gl->w, gl->b, xt, o);
                    wave_lds_sync();   // the tile's inputs are read: its words take their place
                    mixture_words(o, [&](int k, uint32_t sidx, uint32_t v) { mixw[64 * k + sidx] = v; });
                    wave_lds_sync();
                }
            }
        }
        bool to_ray = false;
        ShadowOut so;
        so.emit = false;
        V3 new_o = v3(0.0f), new_d = v3(0.0f), new_thr = v3(0.0f);
        uint32_t new_flags = 0;
        uint4 m = make_uint4(0, 0, 0, 0);
        float cur_ior = 1.0f;
        float scene_bits = 0.0f;
        uint32_t head = kInvalid;          // newest guide record of the path
        float rrs = 1.0f;                  // guided roulette scale of the path
        bool rec_valid = false;            // the continuing path writes a training record
        V3 rec_p = v3(0.0f), rec_n = v3(0.0f), rec_wo = v3(0.0f);
        float rec_pdf = 0.0f, rec_cos = 0.0f;
        V3 add_L = v3(0.0f);
        bool add = false;
        uint32_t pixel = 0;
        if (valid) {
            int q;
            uint32_t j;
            if (e < e1) { q = 0; j = e; }
            else if (e < e2) { q = 1; j = e - e1; }
            else if (e < e3) { q = 2; j = e - e2; }
            else { q = 3; j = e - e3; }
            const uint32_t* qa = q == 0 ? p.q[0] : q == 1 ? p.q[1] : q == 2 ? p.q[2] : p.q[3];
            const uint32_t i = qa[j];
            m = (cur ? p.meta[1] : p.meta[0])[i];
            const uint32_t depth = flags_depth(m.y);
            const bool delta = (m.y & kDeltaBit) != 0u;
            const ViewRecord* vr = p.views + (MULTI ? flags_view(m.y) : 0u);
            const uint32_t scene = vr->scene;
            const SceneDev sc = MULTI ? p.scenes[scene] : sc1;
            const uint32_t limit = vr->depth_limit, rr_start = vr->rr_start;
            const bool view_nee = vr->use_nee != 0u;
            const bool nee = view_nee && sc.light_count > 0u;
            pixel = vr->accum_offset + m.x;
            scene_bits = __uint_as_float(scene);
            const float4 th4 = (cur ? p.thr[1] : p.thr[0])[i];
            const V3 thr = xyz(th4);
            cur_ior = th4.w;
            const float4 hh = p.hit[i];
            const uint32_t prim = __float_as_uint(hh.w);
            if (GUIDED) {
                head = (cur ? p.guide.head[1] : p.guide.head[0])[i];
                rrs = (cur ? p.guide.rr_scale[1] : p.guide.rr_scale[0])[i];
            }
            if (q == 3) {
                // EvaluateEmissiveKernel, MaterialEval.cu:882-912
                if (!view_nee || delta) {
                    const uint32_t mi = reinterpret_cast<const uint32_t*>(sc.shade + prim)[27];
                    const MaterialRecord* mr = sc.materials + mi;
                    add_L = thr * v3(mr->emission[0], mr->emission[1], mr->emission[2]);
                    add = true;
                    if (GUIDED) guide_credit(p.guide.recs, head, add_L);
                }
            } else {
                const float4 o4 = (cur ? p.ray_o[1] : p.ray_o[0])[i];
                const float4 d4 = (cur ? p.ray_d[1] : p.ray_d[0])[i];
                const V3 ro = xyz(o4), rdir = xyz(d4);
                uint32_t mi;
                const Hit h = resolve_hit(sc.shade + prim, ro, rdir, hh.x, hh.y, hh.z, mi);
                const MaterialRecord* mr = sc.materials + mi;
                const float4* mf = reinterpret_cast<const float4*>(mr);
                const float4 m0 = mf[0], m1 = mf[1], m2 = mf[2], m4 = mf[4];
                const V3 nts = xyz(m4);
                const Frame3 fr = shading_frame(h, nts);
                Draws dr = draws_at(fp, vr->seed, ((uint64_t)m.w << 32) | m.z, depth);
                // guide: draw slot 7 picks guide or BSDF; the mixture of this entry
                float u7 = 1.0f;
                const GuideMixRef gm{mixw + (threadIdx.x & 63), 64};
                if (GUIDED && gfrac > 0.0f && q < 2) {
                    // draw slot 7 of the bounce: seven steps ahead in one affine jump
                    Draws d7 = dr;
                    d7.state = jump7.mult * dr.state + jump7.plus * dr.inc;
                    u7 = d7.next();
                }
                const bool by_guide = GUIDED && gfrac > 0.0f && u7 < gfrac;
                float pdf_used = 0.0f;
                float pdf_bsdf = 0.0f;             // the BSDF-sampling density of wi (guided roulette)
                const float u0 = dr.next(), u1 = dr.next(), u2 = dr.next(), u3 = dr.next();
                float l0 = 0.0f, l1 = 0.0f, l2 = 0.0f;
                if (nee) {
                    l0 = dr.next();
                    l1 = dr.next();
                    l2 = dr.next();
                }
                V3 wi = v3(0.0f), weight = v3(0.0f);
                bool scatter = false;
                bool new_delta = false;
                if (q == 0) {
                    // Diffuse, MaterialEval.cu:231-350
                    const V3 albedo = xyz(m0);
                    if (nee) {
                        const LightSample ls = sample_light(sc, l0, l1, l2, h.p);
                        if (!(ls.d2 < 1e-6f)) {
                            const float ci = fmaxf(dot(fr.n, ls.wi), 0.0f);
                            const float gci = fmaxf(dot(h.gn, ls.wi), 0.0f);
                            if (!(ci <= 0.0f || gci <= 0.0f)) {
                                const float cl = fmaxf(dot(ls.gn, v3(0.0f) - ls.wi), 0.0f);
                                if (!(cl <= 0.0f)) {
                                    const V3 brdf = albedo / kPi;
                                    const float G = ci * cl / ls.d2;
                                    const V3 L = ls.emission * thr * brdf * G / ls.pdf;
                                    make_shadow(so, h.p, h.gn, ls.p, L);
                                }
                            }
                        }
                    }
                    wi = by_guide ? guide_sample_folded(gm, h.gn, u2, u0, u1) : fr.from_local(cosine_hemisphere(u0, u1));
                    const float ci = fmaxf(dot(fr.n, wi), 0.0f);
                    const float gci = fmaxf(dot(h.gn, wi), 0.0f);
                    if (!(ci <= 0.0f || gci <= 0.0f)) {
                        if (GUIDED && gfrac > 0.0f) {
                            // one-sample MIS: f cos / (a q + (1-a) cos/pi)
                            const float pdf = gfrac * guide_pdf_folded(gm, wi, h.gn) + (1.0f - gfrac) * ci * kInvPi;
                            if (pdf > 1e-8f) {
                                weight = albedo * (ci * kInvPi / pdf);
                                pdf_used = pdf;
                                pdf_bsdf = ci * kInvPi;
                                scatter = true;
                            }
                        } else {
                            weight = albedo;
                            pdf_used = ci * kInvPi;
                            scatter = true;
                        }
                    }
                } else if (q == 1) {
                    // GGX, MaterialEval.cu:652-880
                    const V3 base = xyz(m0);
                    const float alpha = m1.w;
                    const float metallic = m2.x;
                    const float ior = m2.y;
                    const float df0 = sqr((cur_ior - ior) / (cur_ior + ior));
                    const V3 f0 = lerp3(v3(df0), base, metallic);
                    const V3 diffc = base * (1.0f - metallic);
                    const V3 wo = fr.to_local(v3(0.0f) - rdir);
                    if (nee) {
                        const LightSample ls = sample_light(sc, l0, l1, l2, h.p);
                        if (!(ls.d2 < 1e-6f || ls.pdf <= 0.0f)) {
                            const float gci = dot(h.gn, ls.wi);
                            const float cl = fmaxf(dot(ls.gn, v3(0.0f) - ls.wi), 0.0f);
                            if (!(gci <= 0.0f) && !(cl <= 0.0f)) {
                                const V3 wl = fr.to_local(ls.wi);
                                const float ci = wl.z, co = wo.z;
                                if (!(ci <= 0.0f || co <= 0.0f)) {
                                    const V3 wm = normalize(wo + wl);
                                    if (!(wm.z <= 0.0f)) {
                                        const float owm = dot(wo, wm), iwm = dot(wl, wm);
                                        if (!(owm <= 1e-6f || iwm <= 1e-6f)) {
                                            const V3 F = fresnel_schlick(f0, clampf(owm, 0.0f, 1.0f));
                                            const V3 fd = (v3(1.0f) - F) * diffc * kInvPi;
                                            const V3 fs = F * ggx_d(wm, alpha) * ggx_g(wo, wl, alpha) / (4.0f * ci * co);
                                            const V3 bsdf = fd + fs;
                                            const float G = ci * cl / ls.d2;
                                            const V3 L = ls.emission * thr * bsdf * G / ls.pdf;
                                            make_shadow(so, h.p, h.gn, ls.p, L);
                                        }
                                    }
                                }
                            }
                        }
                    }
                    // EvaluateGgxIndirect, MaterialEval.cu:720-836
                    if (wo.z > 0.0f) {
                        const float dw = luminance(diffc), sw = luminance(f0);
                        const float wsum = dw + sw;
                        if (!(wsum <= 1e-6f)) {
                            const float pd = dw / wsum, ps = sw / wsum;
                            V3 wl;
                            if (by_guide) wl = fr.to_local(guide_sample_folded(gm, h.gn, u2, u0, u1));
                            else if (u2 < pd) wl = cosine_hemisphere(u0, u1);
                            else wl = reflect(-wo, ggx_sample_wm(wo, alpha, u0, u1));
                            const float ci = wl.z, co = wo.z;
                            if (!(ci <= 0.0f)) {
                                const V3 wm = normalize(wl + wo);
                                if (!(wm.z <= 0.0f)) {
                                    const float owm = dot(wo, wm);
                                    if (!(owm <= 1e-6f)) {
                                        const V3 F = fresnel_schlick(f0, clampf(owm, 0.0f, 1.0f));
                                        const V3 fd = (v3(1.0f) - F) * diffc * kInvPi;
                                        const V3 fs = F * ggx_d(wm, alpha) * ggx_g(wo, wl, alpha) / (4.0f * ci * co);
                                        const float pdfd = ci * kInvPi;
                                        const float pdfs = ggx_dv(wo, wm, alpha) / (4.0f * owm);
                                        const V3 ww = fr.from_local(wl);
                                        float pdf = pd * pdfd + ps * pdfs;
                                        pdf_bsdf = pdf;
                                        if (GUIDED && gfrac > 0.0f)
                                            pdf = gfrac * guide_pdf_folded(gm, ww, h.gn) + (1.0f - gfrac) * pdf;
                                        if (!(pdf <= 1e-8f)) {
                                            if (!(dot(ww, h.gn) <= 0.0f)) {
                                                wi = ww;
                                                weight = (fd + fs) * (ci / pdf);
                                                pdf_used = pdf;
                                                scatter = true;
                                            }
                                        }
                                    }
                                }
                            }
                        }
                    }
                } else {
                    // Mirror, MaterialEval.cu:352-392
                    wi = reflect(rdir, fr.n);
                    weight = xyz(m1);
                    scatter = true;
                    new_delta = true;
                }
                if (scatter) {
                    // Guided paths: roulette on the throughput BSDF sampling would
                    // have carried (carried throughput x prod of mixture pdf / BSDF
                    // pdf), so a direction the guide favours is not killed for the
                    // lower weight its higher density gives it.
                    if (GUIDED && gfrac > 0.0f && q < 2)
                        rrs = clampf(rrs * (pdf_bsdf > 0.0f ? pdf_used / pdf_bsdf : 1e6f), 1e-6f, 1e6f);
                    const Continuation k = continue_path(depth, thr, limit, rr_start, weight, u3, GUIDED ? rrs : 1.0f);
                    if (!k.terminated) {
                        // SpawnScatteredRay, MaterialEval.cu:148-156
                        new_o = ro + rdir * hh.x + h.gn * 1e-4f;
                        new_d = wi;
                        new_thr = k.thr;
                        new_flags = k.depth | (m.y & 0x00FFFF00u) | (new_delta ? kDeltaBit : 0u);
                        to_ray = true;
                        if (GUIDED && q < 2 && (m.z & rec_mask) == 0u) {
                            rec_valid = true;
                            rec_p = h.p;
                            rec_n = h.gn;
                            rec_wo = v3(0.0f) - rdir;
                            rec_pdf = pdf_used;
                            rec_cos = fabsf(dot(fr.n, wi));
                        }
                    }
                }
            }
        }
        if (add) {
            float* acc = reinterpret_cast<float*>(p.accum + pixel);
            atomicAdd(acc + 0, add_L.x);
            atomicAdd(acc + 1, add_L.y);
            atomicAdd(acc + 2, add_L.z);
        }
        uint32_t ray_pos, sh_pos, rec_pos = 0;
        if (GUIDED)
            block_reserve3(to_ray, so.emit, rec_valid, &c->ray[nxt], &c->shadow[cur], &c->records, ray_pos, sh_pos,
                           rec_pos);
        else
            block_reserve2(to_ray, so.emit, &c->ray[nxt], &c->shadow[cur], ray_pos, sh_pos);
        if (GUIDED) {
            if (so.emit) p.guide.sh_head[sh_pos] = head;
            if (to_ray) {
                // the recording paths' vertices take consecutive record slots
                uint32_t new_head = head;
                if (rec_valid && rec_pos < p.guide.rec_capacity) {
                    float4* rp = reinterpret_cast<float4*>(p.guide.recs + rec_pos);
                    rp[0] = make_float4(rec_p.x, rec_p.y, rec_p.z, __uint_as_float(head));
                    rp[1] = make_float4(rec_n.x, rec_n.y, rec_n.z, rec_pdf);
                    rp[2] = make_float4(rec_wo.x, rec_wo.y, rec_wo.z, rec_cos);
                    rp[3] = make_float4(new_d.x, new_d.y, new_d.z, 0.0f);
                    rp[4] = make_float4(new_thr.x, new_thr.y, new_thr.z, 0.0f);
                    rp[5] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                    new_head = rec_pos;
                }
                (nxt ? p.guide.head[1] : p.guide.head[0])[ray_pos] = new_head;
                (nxt ? p.guide.rr_scale[1] : p.guide.rr_scale[0])[ray_pos] = rrs;
            }
        }
        if (so.emit) {
            p.sh_o[sh_pos] = f4(so.o, so.tmax);
            p.sh_d[sh_pos] = f4(so.d, scene_bits);
            p.sh_L[sh_pos] = f4(so.L, __uint_as_float(pixel));
        }
        if (to_ray) {
            (nxt ? p.ray_o[1] : p.ray_o[0])[ray_pos] = f4(new_o, scene_bits);
            (nxt ? p.ray_d[1] : p.ray_d[0])[ray_pos] = f4(new_d, kMaxT);
            (nxt ? p.thr[1] : p.thr[0])[ray_pos] = f4(new_thr, cur_ior);
            (nxt ? p.meta[1] : p.meta[0])[ray_pos] = make_uint4(m.x, new_flags, m.z, m.w);
        }
    }
}

// Ring slot of the view owning accumulation pixel i among the views with
// virtual index in [lo, hi), whose accum_offset increases with the index.
__device__ __forceinline__ uint32_t view_of_pixel(const ViewRecord* views, uint32_t lo, uint32_t hi, uint32_t i) {
    while (hi - lo > 1) {
        const uint32_t mid = lo + ((hi - lo) >> 1);
        if (views[mid & (kViewRing - 1u)].accum_offset <= i) lo = mid;
        else hi = mid;
    }
    return lo & (kViewRing - 1u);
}

// PostprocessKernel, Image.cu:34-54, over pixels [begin, end) of the views
// with virtual index in [lo, hi).  With `scale`, the accumulation pixels are
// replaced by radiance / spp, ready for a plain copy to the caller.
__global__ void k_postprocess(PathState p, uint32_t lo, uint32_t hi, uint32_t begin, uint32_t end, uint32_t scale,
                              uint32_t* invalid) {
    for (uint32_t i = begin + blockIdx.x * blockDim.x + threadIdx.x; i < end; i += gridDim.x * blockDim.x) {
        const float inv_spp = p.views[view_of_pixel(p.views, lo, hi, i)].inv_spp;
        const float4 a = p.accum[i];
        const bool ok = isfinite(a.x) && a.x >= 0.0f && isfinite(a.y) && a.y >= 0.0f && isfinite(a.z) && a.z >= 0.0f;
        if (!ok) atomicAdd(invalid, 1u);
        const float inv_gamma = 1.0f / 2.2f;
        float c[3] = {a.x * inv_spp, a.y * inv_spp, a.z * inv_spp};
        uint32_t out = 0xFF000000u;
        for (int k = 0; k < 3; k++) {
            const float t = c[k] / (c[k] + 1.0f);
            const float g = powf(t, inv_gamma);
            const uint32_t b = (uint32_t)(unsigned char)clampf(g * 255.0f, 0.0f, 255.0f);
            out |= b << (8 * k);
        }
        p.framebuffer[i] = out;
        if (scale) p.accum[i] = make_float4(c[0], c[1], c[2], a.w * inv_spp);
    }
}

__global__ void k_alias_sample(SceneDev sc, const float* u, uint32_t n, uint32_t* index, float* pmf) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        float pm;
        index[i] = alias_pick(sc.bins, sc.light_count, u[i], pm);
        pmf[i] = pm;
    }
}

// Writes words handed over by value: a stream-ordered update of small device
// tables (views, scene table, frame constants) with no host staging buffer.
struct WordPayload {
    uint32_t w[kWritePayloadWords];
};

__global__ void k_write_words(uint32_t* dst, WordPayload src, uint32_t n) {
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) dst[i] = src.w[i];
}

// The allocation-free trace's scratch (counters + stack spill).  In host code
// `scratch` names a host-side stand-in: only the address that
// hipGetSymbolAddress returns for the current device may reach a kernel or a
// copy.
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

__global__ void k_zero_u32(unsigned int* ptr, uint32_t n) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) ptr[i] = 0u;
}

// ================================================================ launchers

namespace {

uint32_t grid_for(uint32_t n, uint32_t block, uint32_t max_grid) {
    uint64_t g = ((uint64_t)n + block - 1) / block;
    if (g < 1) g = 1;
    if (g > max_grid) g = max_grid;
    return (uint32_t)g;
}

template <int BLOCK, int LDS, int W, bool ANY, bool COUNT, bool MULTI, class Src>
hipError_t run_trace(hipStream_t s, const Src& src, const SceneDev& sc, const SceneDev* scenes, const TraceConfig& cfg) {
    hipLaunchKernelGGL((k_trace<BLOCK, LDS, W, ANY, COUNT, MULTI, Src>), dim3(cfg.blocks), dim3(BLOCK), 0, s, src, sc,
                       scenes, cfg);
    return hipGetLastError();
}

template <int W, bool ANY, bool COUNT, bool MULTI, class Src>
hipError_t dispatch_block(hipStream_t s, const Src& src, const SceneDev& sc, const SceneDev* scenes,
                          const TraceConfig& cfg) {
    if (cfg.block == 64 && cfg.lds_stack == 16) return run_trace<64, 16, W, ANY, COUNT, MULTI>(s, src, sc, scenes, cfg);
    if (cfg.block == 128 && cfg.lds_stack == 16) return run_trace<128, 16, W, ANY, COUNT, MULTI>(s, src, sc, scenes, cfg);
    if (cfg.block == 256 && cfg.lds_stack == 8) return run_trace<256, 8, W, ANY, COUNT, MULTI>(s, src, sc, scenes, cfg);
    if (cfg.block == 512 && cfg.lds_stack == 8) return run_trace<512, 8, W, ANY, COUNT, MULTI>(s, src, sc, scenes, cfg);
    return hipErrorInvalidValue;
}

// Node width from the scene record, block size from the configuration.
template <bool ANY, bool COUNT, bool MULTI, class Src>
hipError_t dispatch_trace(hipStream_t s, const Src& src, const SceneDev& sc, const SceneDev* scenes,
                          const TraceConfig& cfg) {
    if (sc.node_width == 4 && sc.qnodes8 && (ANY || (sc.node_quantized & 8u)))
        return dispatch_block<kNodeBvh8Q, ANY, COUNT, MULTI>(s, src, sc, scenes, cfg);
    if (sc.node_width == 4 && (sc.node_quantized & (ANY ? 2u : 1u)))
        return dispatch_block<kNodeBvh4Q, ANY, COUNT, MULTI>(s, src, sc, scenes, cfg);
    if (sc.node_width == 4) return dispatch_block<kNodeBvh4, ANY, COUNT, MULTI>(s, src, sc, scenes, cfg);
    if (sc.node_width == 2) return dispatch_block<kNodeBvh2, ANY, COUNT, MULTI>(s, src, sc, scenes, cfg);
    return hipErrorInvalidValue;
}

template <int BLOCK, int LDS, int NODE, bool ANY, bool COUNT, bool MULTI, class Src>
int occ_one() {
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_trace<BLOCK, LDS, NODE, ANY, COUNT, MULTI, Src>, BLOCK, 0) !=
        hipSuccess)
        return 0;
    return nb;
}

template <int NODE, bool ANY, bool COUNT, bool MULTI, class Src>
int occ(const TraceConfig& cfg) {
    if (cfg.block == 64 && cfg.lds_stack == 16) return occ_one<64, 16, NODE, ANY, COUNT, MULTI, Src>();
    if (cfg.block == 128 && cfg.lds_stack == 16) return occ_one<128, 16, NODE, ANY, COUNT, MULTI, Src>();
    if (cfg.block == 256 && cfg.lds_stack == 8) return occ_one<256, 8, NODE, ANY, COUNT, MULTI, Src>();
    if (cfg.block == 512 && cfg.lds_stack == 8) return occ_one<512, 8, NODE, ANY, COUNT, MULTI, Src>();
    return 0;
}

template <int NODE, bool MULTI>
int occ_kind(const TraceConfig& cfg, bool any, bool count) {
    if (any) return count ? occ<NODE, true, true, MULTI, WavefrontShadowSrc>(cfg) : occ<NODE, true, false, MULTI, WavefrontShadowSrc>(cfg);
    return count ? occ<NODE, false, true, MULTI, WavefrontClosestSrc>(cfg) : occ<NODE, false, false, MULTI, WavefrontClosestSrc>(cfg);
}

// the 8-wide quantized nodes: any hit, and closest hit where bvh_quantized has bit 3
template <bool MULTI>
int occ_wide8(const TraceConfig& cfg, bool any, bool count) {
    return occ_kind<kNodeBvh8Q, MULTI>(cfg, any, count);
}

}  // namespace

bool trace_config_supported(uint32_t block, uint32_t lds_stack) {
    return (block == 64 && lds_stack == 16) || (block == 128 && lds_stack == 16) || (block == 256 && lds_stack == 8) ||
           (block == 512 && lds_stack == 8);
}

int trace_occupancy(const TraceConfig& cfg, bool any, bool count, int node_kind, bool multi) {
    if (node_kind == 2) return multi ? occ_wide8<true>(cfg, any, count) : occ_wide8<false>(cfg, any, count);
    if (node_kind >= 1) return multi ? occ_kind<kNodeBvh4Q, true>(cfg, any, count) : occ_kind<kNodeBvh4Q, false>(cfg, any, count);
    return multi ? occ_kind<kNodeBvh4, true>(cfg, any, count) : occ_kind<kNodeBvh4, false>(cfg, any, count);
}

hipError_t launch_frame_begin(hipStream_t s, const PathState& p) {
    hipLaunchKernelGGL(k_frame_begin, dim3(1), dim3(64), 0, s, p);
    return hipGetLastError();
}

hipError_t launch_fill_accum(hipStream_t s, const PathState& p, uint32_t begin, uint32_t end) {
    if (end <= begin) return hipSuccess;
    hipLaunchKernelGGL(k_fill_accum, dim3(grid_for(end - begin, 256, 4096)), dim3(256), 0, s, p.accum, begin, end);
    return hipGetLastError();
}

hipError_t launch_generate(hipStream_t s, const PathState& p, bool init, int nxt, uint32_t grid, uint32_t block) {
    hipLaunchKernelGGL(k_generate, dim3(grid), dim3(block), 0, s, p, init ? 1 : 0, nxt);
    return hipGetLastError();
}

hipError_t launch_closest(hipStream_t s, const PathState& p, const SceneDev& sc, int cur, bool count, bool multi,
                          const TraceConfig& cfg) {
    WavefrontClosestSrc src{p, cur};
    if (multi)
        return count ? dispatch_trace<false, true, true>(s, src, sc, p.scenes, cfg)
                     : dispatch_trace<false, false, true>(s, src, sc, p.scenes, cfg);
    return count ? dispatch_trace<false, true, false>(s, src, sc, p.scenes, cfg)
                 : dispatch_trace<false, false, false>(s, src, sc, p.scenes, cfg);
}

hipError_t launch_bin(hipStream_t s, const PathState& p, int cur, uint32_t grid) {
    hipLaunchKernelGGL(k_bin, dim3(grid), dim3(kBinBlock), 0, s, p, cur);
    return hipGetLastError();
}

hipError_t launch_shade(hipStream_t s, const PathState& p, const SceneDev& sc, int cur, bool multi, bool guided,
                        uint32_t grid, uint32_t block) {
    if (block != 256 || (multi && guided)) return hipErrorInvalidValue;
    // the guided kernel holds the guide network in 159 KB of LDS: one block of
    // 16 waves per CU where the others run eight blocks of four
    if (guided)
        hipLaunchKernelGGL((k_shade<false, true>), dim3(std::max(1u, grid / 8u)), dim3(64 * kInferWaves), 0, s, p, sc,
                           cur);
    else if (multi) hipLaunchKernelGGL((k_shade<true, false>), dim3(grid), dim3(256), 0, s, p, sc, cur);
    else hipLaunchKernelGGL((k_shade<false, false>), dim3(grid), dim3(256), 0, s, p, sc, cur);
    return hipGetLastError();
}

hipError_t launch_shadow(hipStream_t s, const PathState& p, const SceneDev& sc, int cur, bool count, bool multi,
                         const TraceConfig& cfg) {
    WavefrontShadowSrc src{p, cur};
    if (multi)
        return count ? dispatch_trace<true, true, true>(s, src, sc, p.scenes, cfg)
                     : dispatch_trace<true, false, true>(s, src, sc, p.scenes, cfg);
    return count ? dispatch_trace<true, true, false>(s, src, sc, p.scenes, cfg)
                 : dispatch_trace<true, false, false>(s, src, sc, p.scenes, cfg);
}

hipError_t launch_prepare(hipStream_t s, const PathState& p, int cur) {
    hipLaunchKernelGGL(k_prepare, dim3(1), dim3(64), 0, s, p, cur);
    return hipGetLastError();
}

hipError_t launch_postprocess(hipStream_t s, const PathState& p, uint32_t lo, uint32_t hi, uint32_t begin,
                              uint32_t end, bool scale, uint32_t* invalid) {
    if (end <= begin || hi <= lo) return hipSuccess;
    if (!invalid) invalid = &p.ctr->invalid_pixels;
    hipLaunchKernelGGL(k_postprocess, dim3(grid_for(end - begin, 256, 4096)), dim3(256), 0, s, p, lo, hi, begin, end,
                       scale ? 1u : 0u, invalid);
    return hipGetLastError();
}

hipError_t launch_trace_batch(hipStream_t s, const SceneDev& sc, const float4* rays, uint32_t n, float4* hits,
                              uint32_t* occluded, uint32_t* nodes, uint32_t* tris, unsigned int* work, bool any,
                              const TraceConfig& cfg) {
    BatchSrc src{rays, hits, occluded, nodes, tris, work, n, nullptr};
    const bool count = nodes != nullptr || tris != nullptr;
    if (any)
        return count ? dispatch_trace<true, true, false>(s, src, sc, nullptr, cfg)
                     : dispatch_trace<true, false, false>(s, src, sc, nullptr, cfg);
    return count ? dispatch_trace<false, true, false>(s, src, sc, nullptr, cfg)
                 : dispatch_trace<false, false, false>(s, src, sc, nullptr, cfg);
}

hipError_t launch_classify(hipStream_t s, const SceneDev& sc, const float4* rays, uint32_t n, uint8_t* qid,
                           uint32_t* const* queues, unsigned int* queue_counts, unsigned int* work,
                           const TraceConfig& cfg, uint32_t bin_grid) {
    BatchSrc src{rays, nullptr, nullptr, nullptr, nullptr, work, n, qid};
    hipError_t e = dispatch_trace<false, false, false>(s, src, sc, nullptr, cfg);
    if (e != hipSuccess) return e;
    QueueSet qs;
    for (int i = 0; i < kQueueSlots; i++) qs.q[i] = queues[i];
    hipLaunchKernelGGL(k_bin_batch, dim3(bin_grid), dim3(kBinBlock), 0, s, qid, n, qs, queue_counts);
    return hipGetLastError();
}

hipError_t launch_alias_sample(hipStream_t s, const SceneDev& sc, const float* u, uint32_t n, uint32_t* index, float* pmf) {
    hipLaunchKernelGGL(k_alias_sample, dim3(grid_for(n, 256, 4096)), dim3(256), 0, s, sc, u, n, index, pmf);
    return hipGetLastError();
}

hipError_t noalloc_scratch_address(void** ptr) {
    constexpr int kMaxDevices = 64;
    static std::mutex mu;
    static void* addr[kMaxDevices] = {};
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(mu);
    if (!addr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        addr[dev] = p;
    }
    *ptr = addr[dev];
    return hipSuccess;
}

// Top-of-tree table of a tree (one thread): breadth-first from the root, each
// internal child that still fits gets the next table slot and its reference
// in the copied parent becomes kTopBit | slot.
template <class Node, int A>
__global__ void k_build_top(const Node* nodes, Node* table) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    constexpr uint32_t kMax = kTopBytes / sizeof(Node);
    uint32_t* gidx = reinterpret_cast<uint32_t*>(table + kMax);
    uint32_t count = 1;
    gidx[0] = 0;
    for (uint32_t i = 0; i < count; i++) {
        Node n = nodes[gidx[i]];
        for (int k = 0; k < A; k++) {
            const uint32_t c = n.child[k];
            if (c == kInvalid || ref_is_leaf(c) || count >= kMax) continue;
            gidx[count] = c;
            n.child[k] = kTopBit | count;
            count++;
        }
        table[i] = n;
    }
    for (uint32_t i = count; i < kMax; i++) {
        table[i] = Node{};
        gidx[i] = 0u;
    }
}

size_t top_table_bytes(int kind) {
    const size_t node = kind == kNodeBvh4 ? 128 : 64;
    return kTopBytes + kTopBytes / node * sizeof(uint32_t);
}

hipError_t launch_build_top(hipStream_t s, int kind, const void* nodes, void* table) {
    switch (kind) {
    case kNodeBvh4:
        hipLaunchKernelGGL((k_build_top<Bvh4Node, 4>), dim3(1), dim3(64), 0, s, static_cast<const Bvh4Node*>(nodes),
                           static_cast<Bvh4Node*>(table));
        break;
    case kNodeBvh2:
        hipLaunchKernelGGL((k_build_top<Bvh2Node, 2>), dim3(1), dim3(64), 0, s, static_cast<const Bvh2Node*>(nodes),
                           static_cast<Bvh2Node*>(table));
        break;
    case kNodeBvh4Q:
        hipLaunchKernelGGL((k_build_top<Bvh4QNode, 4>), dim3(1), dim3(64), 0, s, static_cast<const Bvh4QNode*>(nodes),
                           static_cast<Bvh4QNode*>(table));
        break;
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_write_words(hipStream_t s, void* dst, const void* src, size_t bytes) {
    if (bytes % 4 != 0) return hipErrorInvalidValue;
    const uint32_t* w = static_cast<const uint32_t*>(src);
    uint32_t* d = static_cast<uint32_t*>(dst);
    size_t left = bytes / 4;
    while (left > 0) {
        WordPayload pl;
        const uint32_t n = (uint32_t)(left < kWritePayloadWords ? left : kWritePayloadWords);
        for (uint32_t i = 0; i < n; i++) pl.w[i] = w[i];
        hipLaunchKernelGGL(k_write_words, dim3(1), dim3(256), 0, s, d, pl, n);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
        w += n;
        d += n;
        left -= n;
    }
    return hipSuccess;
}

hipError_t launch_zero_u32(hipStream_t s, unsigned int* ptr, uint32_t n) {
    hipLaunchKernelGGL(k_zero_u32, dim3(grid_for(n, 256, 1024)), dim3(256), 0, s, ptr, n);
    return hipGetLastError();
}

}  // namespace pt
