// GPU BVH4 builder for large scenes: a linear BVH (Karras 2012, "Maximizing
// Parallelism in the Construction of BVHs, Octrees, and k-d Trees") over
// 30-bit Morton codes of the triangle centroids, refitted bottom-up level by
// level, then collapsed top-down into the 128-byte BVH4 nodes the traversal
// kernels read (same child-box padding, leaf references and empty-slot rule as
// the host builder).  It replaces the host's recursive build
// (HostBvh.cpp:22-100, DeviceBvh.cpp:7-52) where the scene is too large for
// it; every traversal result is independent of the tree.
//
// Level-synchronous passes (one launch per tree level) keep every cross-block
// dependency at a kernel boundary, so no device-scope fences are needed.
#include <rocprim/device/device_radix_sort.hpp>

#include <algorithm>
#include <vector>

#include "bvh_gpu.hpp"

namespace pt {

namespace {

constexpr uint32_t kBlock = 256;
constexpr size_t kMaxLevels = 256;   // a 30-bit Morton tree over < 2^27 leaves is far shallower

inline uint32_t blocks_for(uint32_t n) { return std::max(1u, std::min((n + kBlock - 1) / kBlock, 65536u)); }

// ---------------------------------------------------------------- triangle boxes, Morton codes

// Box of triangle t (lo.xyz, hi.xyz) and the per-block union of the centroids.
__global__ void k_tri_boxes(const float* __restrict__ pos, uint32_t n, float4* __restrict__ lo4, float4* __restrict__ hi4,
                            float* __restrict__ cpart) {
    __shared__ float red[6][kBlock];
    float cl[3] = {INFINITY, INFINITY, INFINITY}, ch[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t t = blockIdx.x * kBlock + threadIdx.x; t < n; t += gridDim.x * kBlock) {
        const float* p = pos + 9 * (size_t)t;
        float lo[3], hi[3];
        for (int a = 0; a < 3; a++) {
            lo[a] = fminf(fminf(p[a], p[3 + a]), p[6 + a]);
            hi[a] = fmaxf(fmaxf(p[a], p[3 + a]), p[6 + a]);
            const float c = 0.5f * (lo[a] + hi[a]);
            cl[a] = fminf(cl[a], c);
            ch[a] = fmaxf(ch[a], c);
        }
        lo4[t] = make_float4(lo[0], lo[1], lo[2], 0.0f);
        hi4[t] = make_float4(hi[0], hi[1], hi[2], 0.0f);
    }
    for (int a = 0; a < 3; a++) {
        red[a][threadIdx.x] = cl[a];
        red[3 + a][threadIdx.x] = ch[a];
    }
    __syncthreads();
    for (uint32_t s = kBlock / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s)
            for (int a = 0; a < 3; a++) {
                red[a][threadIdx.x] = fminf(red[a][threadIdx.x], red[a][threadIdx.x + s]);
                red[3 + a][threadIdx.x] = fmaxf(red[3 + a][threadIdx.x], red[3 + a][threadIdx.x + s]);
            }
        __syncthreads();
    }
    if (threadIdx.x < 6) cpart[(size_t)blockIdx.x * 6 + threadIdx.x] = red[threadIdx.x][0];
}

// Union of the per-block centroid bounds (one block).
__global__ void k_reduce_bounds(const float* __restrict__ cpart, uint32_t parts, float* __restrict__ out6) {
    __shared__ float red[6][kBlock];
    float cl[3] = {INFINITY, INFINITY, INFINITY}, ch[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t b = threadIdx.x; b < parts; b += kBlock)
        for (int a = 0; a < 3; a++) {
            cl[a] = fminf(cl[a], cpart[(size_t)b * 6 + a]);
            ch[a] = fmaxf(ch[a], cpart[(size_t)b * 6 + 3 + a]);
        }
    for (int a = 0; a < 3; a++) {
        red[a][threadIdx.x] = cl[a];
        red[3 + a][threadIdx.x] = ch[a];
    }
    __syncthreads();
    for (uint32_t s = kBlock / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s)
            for (int a = 0; a < 3; a++) {
                red[a][threadIdx.x] = fminf(red[a][threadIdx.x], red[a][threadIdx.x + s]);
                red[3 + a][threadIdx.x] = fmaxf(red[3 + a][threadIdx.x], red[3 + a][threadIdx.x + s]);
            }
        __syncthreads();
    }
    if (threadIdx.x < 6) out6[threadIdx.x] = red[threadIdx.x][0];
}

// 10 bits of x spread to every third bit.
__device__ __forceinline__ uint32_t spread3(uint32_t x) {
    x = (x | (x << 16)) & 0x030000FFu;
    x = (x | (x << 8)) & 0x0300F00Fu;
    x = (x | (x << 4)) & 0x030C30C3u;
    x = (x | (x << 2)) & 0x09249249u;
    return x;
}

__global__ void k_morton(const float4* __restrict__ lo4, const float4* __restrict__ hi4, uint32_t n,
                         const float* __restrict__ cb, uint32_t* __restrict__ keys, uint32_t* __restrict__ ids) {
    const float lo[3] = {cb[0], cb[1], cb[2]};
    float inv[3];
    for (int a = 0; a < 3; a++) {
        const float e = cb[3 + a] - cb[a];
        inv[a] = e > 0.0f ? 1.0f / e : 0.0f;
    }
    for (uint32_t t = blockIdx.x * kBlock + threadIdx.x; t < n; t += gridDim.x * kBlock) {
        const float4 l = lo4[t], h = hi4[t];
        const float c[3] = {0.5f * (l.x + h.x), 0.5f * (l.y + h.y), 0.5f * (l.z + h.z)};
        uint32_t q[3];
        for (int a = 0; a < 3; a++) {
            const float u = fminf(fmaxf((c[a] - lo[a]) * inv[a], 0.0f), 1.0f);
            q[a] = min((uint32_t)(u * 1024.0f), 1023u);
        }
        keys[t] = (spread3(q[0]) << 2) | (spread3(q[1]) << 1) | spread3(q[2]);
        ids[t] = t;
    }
}

// ---------------------------------------------------------------- hierarchy (Karras 2012)

// Length of the common prefix of sorted keys i and j, the index breaking ties
// between equal keys; -1 outside [0, n).
__device__ __forceinline__ int delta(const uint32_t* keys, uint32_t n, int i, int j) {
    if (j < 0 || j >= (int)n) return -1;
    const uint32_t a = keys[i], b = keys[j];
    if (a == b) return 32 + __clz((uint32_t)i ^ (uint32_t)j);
    return __clz(a ^ b);
}

// Internal node i of n - 1: its child references (leaf k = kLeafBit | k) and
// its range [first, last] of sorted leaves.
__global__ void k_karras(const uint32_t* __restrict__ keys, uint32_t n, uint2* __restrict__ child,
                         uint2* __restrict__ range) {
    for (uint32_t ui = blockIdx.x * kBlock + threadIdx.x; ui + 1 < n; ui += gridDim.x * kBlock) {
        const int i = (int)ui;
        const int d = delta(keys, n, i, i + 1) - delta(keys, n, i, i - 1) >= 0 ? 1 : -1;
        const int dmin = delta(keys, n, i, i - d);
        int lmax = 2;
        while (delta(keys, n, i, i + lmax * d) > dmin) lmax *= 2;
        int l = 0;
        for (int t = lmax / 2; t >= 1; t /= 2)
            if (delta(keys, n, i, i + (l + t) * d) > dmin) l += t;
        const int j = i + l * d;
        const int dnode = delta(keys, n, i, j);
        int s = 0;
        for (int t = (l + 1) / 2;; t = (t + 1) / 2) {
            if (delta(keys, n, i, i + (s + t) * d) > dnode) s += t;
            if (t == 1) break;
        }
        const int gamma = i + s * d + min(d, 0);
        const int first = min(i, j), last = max(i, j);
        const uint32_t left = first == gamma ? (kLeafBit | (uint32_t)gamma) : (uint32_t)gamma;
        const uint32_t right = last == gamma + 1 ? (kLeafBit | (uint32_t)(gamma + 1)) : (uint32_t)(gamma + 1);
        child[i] = make_uint2(left, right);
        range[i] = make_uint2((uint32_t)first, (uint32_t)last);
    }
}

// One breadth-first level: the internal children of the level's nodes form
// the next level, appended after it (next_count counts them; writes past
// `room` are dropped and show as a count above it).
__global__ void k_level(const uint2* __restrict__ child, const uint32_t* __restrict__ cur, uint32_t count,
                        uint32_t* __restrict__ nxt, uint32_t room, unsigned int* __restrict__ next_count) {
    for (uint32_t k = blockIdx.x * kBlock + threadIdx.x; k < count; k += gridDim.x * kBlock) {
        const uint2 c = child[cur[k]];
        if (!(c.x & kLeafBit)) {
            const uint32_t at = atomicAdd(next_count, 1u);
            if (at < room) nxt[at] = c.x;
        }
        if (!(c.y & kLeafBit)) {
            const uint32_t at = atomicAdd(next_count, 1u);
            if (at < room) nxt[at] = c.y;
        }
    }
}

__device__ __forceinline__ void child_box(uint32_t ref, const float4* blo, const float4* bhi, const float4* tlo,
                                          const float4* thi, const uint32_t* ids, float4& lo, float4& hi) {
    if (ref & kLeafBit) {
        const uint32_t t = ids[ref & ~kLeafBit];
        lo = tlo[t];
        hi = thi[t];
    } else {
        lo = blo[ref];
        hi = bhi[ref];
    }
}

// Bounds of the nodes of one level from their children (the deeper level is done).
__global__ void k_refit(const uint2* __restrict__ child, const uint32_t* __restrict__ nodes, uint32_t count,
                        const float4* __restrict__ tlo, const float4* __restrict__ thi, const uint32_t* __restrict__ ids,
                        float4* __restrict__ blo, float4* __restrict__ bhi) {
    for (uint32_t k = blockIdx.x * kBlock + threadIdx.x; k < count; k += gridDim.x * kBlock) {
        const uint32_t i = nodes[k];
        const uint2 c = child[i];
        float4 al, ah, bl, bh;
        child_box(c.x, blo, bhi, tlo, thi, ids, al, ah);
        child_box(c.y, blo, bhi, tlo, thi, ids, bl, bh);
        blo[i] = make_float4(fminf(al.x, bl.x), fminf(al.y, bl.y), fminf(al.z, bl.z), 0.0f);
        bhi[i] = make_float4(fmaxf(ah.x, bh.x), fmaxf(ah.y, bh.y), fmaxf(ah.z, bh.z), 0.0f);
    }
}

// ---------------------------------------------------------------- collapse to BVH4

struct CollapseCtx {
    const uint2* child;
    const uint2* range;
    const float4* blo;
    const float4* bhi;
    const float4* tlo;
    const float4* thi;
    const uint32_t* ids;
    Bvh4Node* nodes;
    unsigned int* node_count;
    unsigned int* stats;      // [0] leaves, [1] largest leaf, [2] error flag (node capacity)
    uint32_t node_capacity;
    uint32_t max_leaf;
};

__device__ __forceinline__ uint32_t ref_size(const CollapseCtx& c, uint32_t ref) {
    if (ref & kLeafBit) return 1u;
    const uint2 r = c.range[ref];
    return r.y - r.x + 1u;
}
__device__ __forceinline__ uint32_t ref_first(const CollapseCtx& c, uint32_t ref) {
    return (ref & kLeafBit) ? (ref & ~kLeafBit) : c.range[ref].x;
}
__device__ __forceinline__ float half_area(float4 lo, float4 hi) {
    const float x = hi.x - lo.x, y = hi.y - lo.y, z = hi.z - lo.z;
    return x * y + y * z + z * x;
}

// The host builder's child-box padding (bvh_build.cpp pad_bounds).
__device__ __forceinline__ void pad_box(float lo[3], float hi[3]) {
    const float ex = hi[0] - lo[0], ey = hi[1] - lo[1], ez = hi[2] - lo[2];
    const float largest = fmaxf(ex, fmaxf(ey, ez));
    const float padding = fmaxf(1e-5f * largest, 1e-5f);
    for (int a = 0; a < 3; a++) {
        lo[a] = lo[a] - padding;
        hi[a] = hi[a] + padding;
    }
}

// One level of BVH4 nodes: task k turns BVH2 node src[k] into BVH4 node dst[k],
// expanding the largest-area child that is too big for a leaf until the node
// has four children; children that stay internal become the next level's tasks.
__global__ void k_collapse(CollapseCtx c, const uint2* __restrict__ tasks, uint32_t count, uint2* __restrict__ next,
                           unsigned int* __restrict__ next_count) {
    for (uint32_t k = blockIdx.x * kBlock + threadIdx.x; k < count; k += gridDim.x * kBlock) {
        const uint2 task = tasks[k];
        uint32_t ch[4];
        int n = 0;
        if (task.x & kLeafBit) {
            ch[n++] = task.x;                // a one-triangle scene
        } else {
            const uint2 cc = c.child[task.x];
            ch[n++] = cc.x;
            ch[n++] = cc.y;
        }
        while (n < 4) {
            int pick = -1;
            float best = -1.0f;
            for (int i = 0; i < n; i++) {
                if ((ch[i] & kLeafBit) || ref_size(c, ch[i]) <= c.max_leaf) continue;
                const float a = half_area(c.blo[ch[i]], c.bhi[ch[i]]);
                if (a > best) {
                    best = a;
                    pick = i;
                }
            }
            if (pick < 0) break;
            const uint2 cc = c.child[ch[pick]];
            ch[pick] = cc.x;
            ch[n++] = cc.y;
        }
        Bvh4Node node;
        for (int s = 0; s < 4; s++) {
            node.lox[s] = node.loy[s] = node.loz[s] = INFINITY;
            node.hix[s] = node.hiy[s] = node.hiz[s] = INFINITY;
            node.child[s] = kInvalid;
            node.info[s] = 0;
        }
        for (int s = 0; s < n; s++) {
            float4 l4, h4;
            child_box(ch[s], c.blo, c.bhi, c.tlo, c.thi, c.ids, l4, h4);
            float lo[3] = {l4.x, l4.y, l4.z}, hi[3] = {h4.x, h4.y, h4.z};
            pad_box(lo, hi);
            node.lox[s] = lo[0];
            node.loy[s] = lo[1];
            node.loz[s] = lo[2];
            node.hix[s] = hi[0];
            node.hiy[s] = hi[1];
            node.hiz[s] = hi[2];
            const uint32_t size = ref_size(c, ch[s]);
            if ((ch[s] & kLeafBit) || size <= c.max_leaf) {
                node.child[s] = make_leaf_ref(ref_first(c, ch[s]), size);
                node.info[s] = size;
                atomicAdd(&c.stats[0], 1u);
                atomicMax(&c.stats[1], size);
            } else {
                const uint32_t idx = atomicAdd(c.node_count, 1u);
                if (idx >= c.node_capacity) {
                    atomicOr(&c.stats[2], 1u);
                    continue;
                }
                node.child[s] = idx;
                next[atomicAdd(next_count, 1u)] = make_uint2(ch[s], idx);
            }
        }
        if (task.y < c.node_capacity) c.nodes[task.y] = node;
    }
}

// Triangle records in BVH order (BuildTriangles, Renderer.cpp:159-213).
__global__ void k_tri_records(const float* __restrict__ pos, const uint32_t* __restrict__ material_ids,
                              const MaterialRecord* __restrict__ materials, const uint32_t* __restrict__ ids,
                              uint32_t n, TriRecord* __restrict__ out) {
    for (uint32_t k = blockIdx.x * kBlock + threadIdx.x; k < n; k += gridDim.x * kBlock) {
        const uint32_t t = ids[k];
        const float* p = pos + 9 * (size_t)t;
        TriRecord r;
        for (int a = 0; a < 3; a++) {
            r.v0[a] = p[a];
            r.e1[a] = p[3 + a] - p[a];
            r.e2[a] = p[6 + a] - p[a];
        }
        r.prim = t;
        r.material = material_ids[t];
        r.queue = materials[r.material].queue;
        out[k] = r;
    }
}

struct Temp {
    std::vector<void*> ptrs;
    ~Temp() {
        for (void* p : ptrs) (void)hipFree(p);
    }
    template <class T>
    hipError_t alloc(T*& out, size_t count) {
        void* p = nullptr;
        const hipError_t e = hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(T));
        if (e != hipSuccess) return e;
        ptrs.push_back(p);
        out = static_cast<T*>(p);
        return hipSuccess;
    }
};

#define GB_TRY(x)                              \
    do {                                       \
        const hipError_t e_ = (x);             \
        if (e_ != hipSuccess) return e_;       \
    } while (0)

}  // namespace

hipError_t gpu_build_bvh4(hipStream_t s, const float* d_positions, const uint32_t* d_material_ids,
                          const MaterialRecord* d_materials, uint32_t n, uint32_t max_leaf, Bvh4Node* d_nodes,
                          uint32_t node_capacity, TriRecord* d_tris, uint32_t* d_order, GpuBvhStats& stats) {
    stats = GpuBvhStats{};
    if (n == 0 || max_leaf < 1 || max_leaf > kMaxLeafSize || n > kMaxLeafFirst) return hipErrorInvalidValue;
    if (node_capacity < gpu_bvh4_node_bound(n)) return hipErrorInvalidValue;
    Temp tmp;
    float4 *tlo, *thi, *blo, *bhi;
    float *cpart, *cb;
    uint32_t *keys, *keys2, *ids;
    uint2 *child, *range;
    uint32_t* order2;
    unsigned int* counters;
    const uint32_t nb = blocks_for(n);
    const uint32_t ni = n > 1 ? n - 1 : 1;
    GB_TRY(tmp.alloc(tlo, n));
    GB_TRY(tmp.alloc(thi, n));
    GB_TRY(tmp.alloc(cpart, (size_t)nb * 6));
    GB_TRY(tmp.alloc(cb, 6));
    GB_TRY(tmp.alloc(keys, n));
    GB_TRY(tmp.alloc(keys2, n));
    GB_TRY(tmp.alloc(ids, n));
    GB_TRY(tmp.alloc(child, ni));
    GB_TRY(tmp.alloc(range, ni));
    GB_TRY(tmp.alloc(blo, ni));
    GB_TRY(tmp.alloc(bhi, ni));
    GB_TRY(tmp.alloc(order2, ni));
    GB_TRY(tmp.alloc(counters, 8));

    hipLaunchKernelGGL(k_tri_boxes, dim3(nb), dim3(kBlock), 0, s, d_positions, n, tlo, thi, cpart);
    hipLaunchKernelGGL(k_reduce_bounds, dim3(1), dim3(kBlock), 0, s, cpart, nb, cb);
    hipLaunchKernelGGL(k_morton, dim3(nb), dim3(kBlock), 0, s, tlo, thi, n, cb, keys, ids);
    GB_TRY(hipGetLastError());

    // sort (key, triangle) pairs by key
    size_t temp_bytes = 0;
    GB_TRY(rocprim::radix_sort_pairs(nullptr, temp_bytes, keys, keys2, ids, d_order, n, 0, 30, s));
    uint8_t* sort_tmp = nullptr;
    GB_TRY(tmp.alloc(sort_tmp, temp_bytes));
    GB_TRY(rocprim::radix_sort_pairs(sort_tmp, temp_bytes, keys, keys2, ids, d_order, n, 0, 30, s));

    // hierarchy and its breadth-first levels (level sizes read back per level)
    std::vector<uint32_t> level_off{0};
    if (n > 1) {
        hipLaunchKernelGGL(k_karras, dim3(blocks_for(n - 1)), dim3(kBlock), 0, s, keys2, n, child, range);
        const uint32_t root = 0;
        GB_TRY(hipMemcpyAsync(order2, &root, sizeof(uint32_t), hipMemcpyHostToDevice, s));
        level_off.push_back(1);
        while (true) {
            const uint32_t b = level_off[level_off.size() - 2], e = level_off.back();
            GB_TRY(hipMemsetAsync(counters, 0, sizeof(unsigned int), s));
            hipLaunchKernelGGL(k_level, dim3(blocks_for(e - b)), dim3(kBlock), 0, s, child, order2 + b, e - b,
                               order2 + e, (n - 1) - e, counters);
            uint32_t got = 0;
            GB_TRY(hipMemcpyAsync(&got, counters, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
            GB_TRY(hipStreamSynchronize(s));
            if (got == 0) break;
            if (got > (n - 1) - e || level_off.size() > kMaxLevels) return hipErrorUnknown;   // not a tree
            level_off.push_back(e + got);
        }
        // bounds bottom-up, deepest level first
        for (size_t L = level_off.size() - 1; L-- > 0;) {
            const uint32_t b = level_off[L], e = level_off[L + 1];
            hipLaunchKernelGGL(k_refit, dim3(blocks_for(e - b)), dim3(kBlock), 0, s, child, order2 + b, e - b, tlo,
                               thi, d_order, blo, bhi);
        }
        GB_TRY(hipGetLastError());
    }

    // collapse, one BVH4 level per launch; tasks ping-pong between two arrays
    uint2 *ta, *tb;
    GB_TRY(tmp.alloc(ta, node_capacity));
    GB_TRY(tmp.alloc(tb, node_capacity));
    GB_TRY(hipMemsetAsync(counters, 0, 8 * sizeof(unsigned int), s));
    const uint2 root_task = make_uint2(n > 1 ? 0u : (kLeafBit | 0u), 0u);
    GB_TRY(hipMemcpyAsync(ta, &root_task, sizeof(uint2), hipMemcpyHostToDevice, s));
    const uint32_t one = 1;
    GB_TRY(hipMemcpyAsync(counters + 0, &one, sizeof(uint32_t), hipMemcpyHostToDevice, s));   // nodes: the root
    CollapseCtx ctx{child, range, blo, bhi, tlo, thi, d_order, d_nodes, counters + 0, counters + 4, node_capacity,
                    max_leaf};
    uint32_t count = 1, levels = 0;
    while (count > 0) {
        if (levels >= kMaxLevels) return hipErrorUnknown;
        GB_TRY(hipMemsetAsync(counters + 1, 0, sizeof(unsigned int), s));
        hipLaunchKernelGGL(k_collapse, dim3(blocks_for(count)), dim3(kBlock), 0, s, ctx, ta, count, tb, counters + 1);
        GB_TRY(hipGetLastError());
        GB_TRY(hipMemcpyAsync(&count, counters + 1, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
        GB_TRY(hipStreamSynchronize(s));
        levels++;
        std::swap(ta, tb);
    }
    hipLaunchKernelGGL(k_tri_records, dim3(nb), dim3(kBlock), 0, s, d_positions, d_material_ids, d_materials, d_order, n,
                       d_tris);
    GB_TRY(hipGetLastError());
    uint32_t h[8];
    GB_TRY(hipMemcpyAsync(h, counters, sizeof(h), hipMemcpyDeviceToHost, s));
    GB_TRY(hipStreamSynchronize(s));
    if (h[6] != 0 || h[0] > node_capacity) return hipErrorUnknown;
    stats.node_count = h[0];
    stats.leaf_count = h[4];
    stats.max_leaf = h[5];
    stats.depth = levels;
    stats.bvh2_levels = (uint32_t)level_off.size() - 1;
    return hipSuccess;
}

__global__ void k_quantize_bvh4(const Bvh4Node* nodes, uint32_t n, Bvh4QNode* out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Bvh4QNode q;
    quantize_bvh4_node(nodes[i], q);
    out[i] = q;
}

hipError_t gpu_quantize_bvh4(hipStream_t s, const Bvh4Node* d_nodes, uint32_t n, Bvh4QNode* d_out) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_quantize_bvh4, dim3((n + 255) / 256), dim3(256), 0, s, d_nodes, n, d_out);
    return hipGetLastError();
}

}  // namespace pt
