// Kernel launchers of the wavefront core (host side), the counterpart of the
// reference's free-function launcher layer
// (Core/Include/Core/Graphics/Cuda/PathTracing/Kernels/Launchers.hpp:18-146).
// Views are POD structs of device pointers passed by value.
#pragma once

#include "guide.hpp"
#include "pt_device.hpp"

namespace pt {

// Queue slots: 0..3 material queues (diffuse, ggx, mirror, emissive), 4 regen.
constexpr int kQueueRegen = 4;
constexpr int kQueueSlots = 5;

// Node kinds of the traversal kernels: fp32 BVH4 (128 B), fp32 BVH2 (64 B),
// quantized BVH4 (64 B) and quantized BVH8 (128 B).
constexpr int kNodeBvh2 = 2;
constexpr int kNodeBvh4 = 4;
constexpr int kNodeBvh4Q = 5;
constexpr int kNodeBvh8Q = 8;

struct SceneDev {
    const void* nodes;              // Bvh4Node[] or Bvh2Node[] by node_width
    const void* qnodes;             // Bvh4QNode[]: the same 4-wide tree quantized, or null
    const void* qnodes8;            // Bvh8QNode[]: the same binary tree 8-wide and quantized (any hit), or null
    const TriRecord* tris;          // BVH order
    const TriRecord* tris8;         // order of the 8-wide tree's leaves (with qnodes8)
    const ShadeRecord* shade;       // original order
    const MaterialRecord* materials;
    const LightRecord* lights;
    const AliasBin* bins;
    uint32_t light_count;
    uint32_t tri_count;
    uint32_t node_width;            // 4 or 2
    uint32_t node_quantized;        // traversals that use qnodes: bit 0 closest hit, bit 1 any hit
    uint32_t node8_count;           // nodes of qnodes8 (breadth first: its first ones are the top of the tree)
    const void* top[2];             // top-of-tree tables of nodes and qnodes (null: none)
};

// Writes the top-of-tree table of `nodes` (kind: kNodeBvh4, kNodeBvh2,
// kNodeBvh4Q or kNodeBvh8Q) into `table` (kTopBytes of nodes, then their
// global indices): one thread, a breadth-first walk from the root.
hipError_t launch_build_top(hipStream_t s, int kind, const void* nodes, void* table);
// Bytes of a top-of-tree table for node kind `kind`.
size_t top_table_bytes(int kind);

// Device state of the guide network and of its training records.
struct GuideDev {
    float* w32;                     // fp32 masters: [layer] { W[out][in], b[out] }
    uint16_t* w;                    // bf16 [layer][out][in]
    uint16_t* wt;                   // bf16 [layer][in][out]
    float* grad;
    float* m;                       // Adam moments
    float* v;
    uint16_t* wf;                   // bf16 weights in the inference kernel's fragment order
    float* bf;                      // biases in the inference kernel's output order
    GuideRecord* recs;              // training records of the current run, Counters::records of them
    uint32_t rec_capacity;
    uint32_t* head[2];              // per path position: newest record of the path, kInvalid if none
    uint32_t* sh_head;              // per shadow entry: newest record before its vertex
    float* rr_scale[2];             // per path position: BSDF-sampling throughput / carried throughput
    GuideConfig cfg;
};

// Path state, compacted every bounce: ray queue k holds, at position i, the
// full state of one live path.  Shading writes each continuing path into the
// next queue; new camera paths fill the free tail.  No slot indirection.
struct PathState {
    float4* ray_o[2];               // origin, scene id (bits); tMin is kMinT
    float4* ray_d[2];               // direction, tMax
    float4* thr[2];                 // throughput, medium ior
    uint4* meta[2];                 // view pixel, flags (depth|view slot|delta), local sample lo, hi
    float4* hit;                    // t, u, v, prim (bits) of current-queue entry i
    uint8_t* qid;                   // target queue of current-queue entry i
    uint32_t* q[kQueueSlots];       // material queues: positions in the current ray queue
    float4* sh_o;                   // shadow queue: origin, tMax
    float4* sh_d;                   // direction, scene id (bits)
    float4* sh_L;                   // radiance, accumulation pixel (bits)
    Counters* ctr;
    const FrameParams* fp;
    const ViewRecord* views;        // view ring, kViewRing records
    const SceneDev* scenes;         // resident scenes by id, kMaxScenes records
    float4* accum;                  // accumulation pixels of all live views
    uint32_t* framebuffer;          // RGBA8, same indexing as accum
    uint32_t capacity;              // path slots
    uint32_t pixels;                // accumulation pixels allocated
    GuideDev guide;                 // head[0] == nullptr: no guide records
};

struct TraceConfig {
    uint32_t block = 128;           // threads per block
    uint32_t lds_stack = 16;        // stack entries per lane held in LDS
    uint32_t blocks = 0;            // persistent grid size
    uint32_t refill = 16;           // idle lanes that trigger a refill from the wave's chunk
    uint32_t chunk = 256;           // rays claimed per wave with one atomic
    uint32_t spill_depth = 0;       // stack entries per lane in the global spill area
    uint2* spill = nullptr;         // [spill_depth][blocks*block]
    uint32_t stack_cap = 0;         // stack entries per lane, at most lds_stack + spill_depth
    unsigned int* overflow = nullptr; // set when a push would exceed stack_cap
};

// Blocks of `block` threads that fit on one CU for the traversal kernels.
// node_kind: 0 fp32 BVH4, 1 quantized BVH4, 2 quantized BVH8 (any hit only)
int trace_occupancy(const TraceConfig& cfg, bool any, bool count, int node_kind, bool multi);
bool trace_config_supported(uint32_t block, uint32_t lds_stack);

hipError_t launch_frame_begin(hipStream_t s, const PathState& p);
// Zeroes accumulation pixels [begin, end).
hipError_t launch_fill_accum(hipStream_t s, const PathState& p, uint32_t begin, uint32_t end);
hipError_t launch_generate(hipStream_t s, const PathState& p, bool init, int nxt, uint32_t grid, uint32_t block);
// `multi`: rays of several scenes (per-ray node/triangle pointers from p.scenes,
// all scenes of node width sc.node_width); otherwise every ray belongs to scene
// `sc`.
hipError_t launch_closest(hipStream_t s, const PathState& p, const SceneDev& sc, int cur, bool count, bool multi,
                          const TraceConfig& cfg);
hipError_t launch_bin(hipStream_t s, const PathState& p, int cur, uint32_t grid);
// `guided`: one-sample MIS of the guide mixture and the BSDF, with training
// records (single view only).
hipError_t launch_shade(hipStream_t s, const PathState& p, const SceneDev& sc, int cur, bool multi, bool guided,
                        uint32_t grid, uint32_t block);
// Any-hit traversal of the shadow rays that the shade of iteration parity `cur` pushed.
hipError_t launch_shadow(hipStream_t s, const PathState& p, const SceneDev& sc, int cur, bool count, bool multi,
                         const TraceConfig& cfg);
hipError_t launch_prepare(hipStream_t s, const PathState& p, int cur);
// Tone-maps pixels [begin, end) of the views with virtual index in [lo, hi);
// with `scale` also divides the accumulation pixels by spp in place.  Pixels
// with non-finite or negative radiance are counted in *invalid (default: the
// frame counters' invalid_pixels).
hipError_t launch_postprocess(hipStream_t s, const PathState& p, uint32_t lo, uint32_t hi, uint32_t begin,
                              uint32_t end, bool scale = false, uint32_t* invalid = nullptr);

// Batch hooks.  rays: n x 2 float4 (origin|tMin, direction|tMax).
// Closest: hits[i] in pt_hit layout (prim bits, t, u, v); any: occluded[i].  nodes/tris may be null.
hipError_t launch_trace_batch(hipStream_t s, const SceneDev& sc, const float4* rays, uint32_t n,
                              float4* hits, uint32_t* occluded, uint32_t* nodes, uint32_t* tris,
                              unsigned int* work, bool any, const TraceConfig& cfg);
// Closest hit + the binning pass of the wavefront: ray indices scattered into 5 queues.
hipError_t launch_classify(hipStream_t s, const SceneDev& sc, const float4* rays, uint32_t n, uint8_t* qid,
                           uint32_t* const* queues, unsigned int* queue_counts, unsigned int* work,
                           const TraceConfig& cfg, uint32_t bin_grid);
hipError_t launch_alias_sample(hipStream_t s, const SceneDev& sc, const float* u, uint32_t n,
                               uint32_t* index, float* pmf);
// Guide network (guide.hip).
hipError_t launch_guide_sync(hipStream_t s, const GuideDev& g);
hipError_t launch_guide_forward(hipStream_t s, const GuideDev& g, const float* features, uint32_t n, float* out);
hipError_t launch_guide_mixture(hipStream_t s, const GuideDev& g, const float* features, uint32_t n, uint32_t* words);
// Blocks (64 records each) of a training step: the number of partial gradients.
inline uint32_t guide_parts(uint32_t batch) { return (batch + 63) / 64; }
hipError_t launch_guide_train(hipStream_t s, const GuideDev& g, const GuideRecord* recs, uint32_t count,
                              uint32_t batch, uint32_t seed, float w_max, float* part, float* loss);
hipError_t launch_guide_sum(hipStream_t s, const float* part, uint32_t parts, float* grad);
hipError_t launch_guide_adam(hipStream_t s, const GuideDev& g, const float* part, uint32_t parts, float lr,
                             uint32_t step, float grad_scale);

// Stream-ordered copy of `bytes` (a multiple of 4) from host memory to device
// memory, passed as kernel arguments: the host buffer may be reused at once.
constexpr uint32_t kWritePayloadWords = 896;
hipError_t launch_write_words(hipStream_t s, void* dst, const void* src, size_t bytes);
hipError_t launch_zero_u32(hipStream_t s, unsigned int* p, uint32_t n);

// Scratch of the allocation-free closest-hit trace (pt_trace_closest_device_noalloc):
// one file-scope device array per GPU, [0, 256) work counter and overflow flag,
// the rest the traversal's global stack spill area.
constexpr size_t kNoAllocScratchBytes = size_t(32) << 20;
constexpr size_t kNoAllocSpillOffset = 256;
// Device address of that array on the current device, looked up once per device.
hipError_t noalloc_scratch_address(void** ptr);

}  // namespace pt
