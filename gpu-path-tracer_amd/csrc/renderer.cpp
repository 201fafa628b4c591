#include "renderer.hpp"
#include "image_io.hpp"
#include "bvh_gpu.hpp"
#include "server.hpp"

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <string>

namespace pt {

// Limits of PathTracerDefaults.hpp:9-35
// path slots: automatic by default (sized to each run, see ensure_pool); the
// reference uses 2^20 (PathTracerDefaults.hpp:9)
constexpr uint32_t kDefaultPool = 0;
constexpr uint32_t kAutoPoolMin = 1u << 22, kAutoPoolMax = 1u << 27;
constexpr uint32_t kMinWidth = 1, kMaxWidth = 7680, kMinHeight = 1, kMaxHeight = 4320;
constexpr uint32_t kMinSpp = 1, kMaxSpp = 10000000;
constexpr uint32_t kMinDepth = 1, kMaxDepth = 100;
constexpr uint32_t kDefaultRrStart = 3;

hipError_t DeviceBuffer::allocate(size_t bytes) {
    if (bytes == 0) bytes = 16;
    if (ptr_ && bytes_ >= bytes) return hipSuccess;
    release();
    hipError_t e = hipMalloc(&ptr_, bytes);
    if (e != hipSuccess) {
        ptr_ = nullptr;
        return e;
    }
    bytes_ = bytes;
    return hipSuccess;
}

void DeviceBuffer::release() {
    if (ptr_) (void)hipFree(ptr_);
    ptr_ = nullptr;
    bytes_ = 0;
}

Renderer::Renderer() {
    tuning_.trace_block = 256;
    tuning_.trace_blocks_per_cu = 0;   // 0 = occupancy
    tuning_.shadow_blocks_per_cu = 0;
    tuning_.shade_block = 256;
    tuning_.pool_size = kDefaultPool;
    tuning_.use_graph = 1;
    tuning_.bvh_max_leaf = 8;
    tuning_.bvh_traversal_cost = 1.0f;
    tuning_.trace_refill = 16;
    tuning_.bvh_width = 4;
    tuning_.bvh_builder = PT_BVH_BUILDER_HOST;
    tuning_.bvh_quantized = 14;         // extension and shadow rays on the compressed 8-wide tree (where the
                                        // tree is built on the device: shadow rays on 64-byte 4-wide quantized
                                        // nodes, extension rays on the fp32 ones)
}

Renderer::~Renderer() {
    // a frame lock still held by the destroying thread is released first, so
    // that the render thread can finish its last publish
    if (frame_locked_here()) (void)UnlockFrame();
    if (server_) (void)ServerStop(nullptr);
    (void)StopRendering();
    invalidate_graph();
    for (auto& e : ring_events_)
        if (e) (void)hipEventDestroy(e);
    for (int i = 0; i < 2; i++) {
        if (fork_events_[i]) (void)hipEventDestroy(fork_events_[i]);
        if (join_events_[i]) (void)hipEventDestroy(join_events_[i]);
    }
    if (h_fp_) (void)hipHostFree(h_fp_);
    if (h_ctr_ring_) (void)hipHostFree(h_ctr_ring_);
    if (h_req_invalid_) (void)hipHostFree(h_req_invalid_);
    for (uint32_t* b : h_frame_)
        if (b) (void)hipHostFree(b);
    if (side_stream_) (void)hipStreamDestroy(side_stream_);
    if (stream_) (void)hipStreamDestroy(stream_);
}

pt_status Renderer::check_hip(hipError_t e, const char* what) {
    if (e == hipSuccess) return PT_OK;
    const pt_status s = (e == hipErrorOutOfMemory) ? PT_ERROR_OUT_OF_MEMORY : PT_ERROR_DEVICE;
    return fail(s, std::string(what) + ": " + hipGetErrorString(e));
}


pt_status Renderer::init() {
    int count = 0;
    PT_TRY_HIP(hipGetDeviceCount(&count), "hipGetDeviceCount");
    if (count < 1) return fail(PT_ERROR_DEVICE, "no HIP device is visible");
    PT_TRY_HIP(hipGetDevice(&device_), "hipGetDevice");
    PT_TRY_HIP(hipDeviceGetAttribute(&cu_count_, hipDeviceAttributeMultiprocessorCount, device_), "CU count");
    PT_TRY_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), "hipStreamCreate");
    PT_TRY_HIP(hipStreamCreateWithFlags(&side_stream_, hipStreamNonBlocking), "hipStreamCreate");
    for (int i = 0; i < 2; i++) {
        PT_TRY_HIP(hipEventCreateWithFlags(&fork_events_[i], hipEventDisableTiming), "hipEventCreate");
        PT_TRY_HIP(hipEventCreateWithFlags(&join_events_[i], hipEventDisableTiming), "hipEventCreate");
    }
    PT_TRY_HIP(hipHostMalloc((void**)&h_fp_, sizeof(FrameParams), hipHostMallocDefault), "hipHostMalloc");
    PT_TRY_HIP(hipHostMalloc((void**)&h_ctr_ring_, sizeof(Counters) * kRing, hipHostMallocDefault), "hipHostMalloc");
    for (auto& e : ring_events_) PT_TRY_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate");
    PT_TRY_HIP(d_ctr_.allocate(sizeof(Counters)), "counters");
    PT_TRY_HIP(d_fp_.allocate(sizeof(FrameParams)), "frame params");
    PT_TRY_HIP(d_work_.allocate(64), "work counter");
    PT_TRY_HIP(d_overflow_.allocate(64), "overflow flag");
    PT_TRY_HIP(d_scene_table_.allocate(sizeof(SceneDev) * kMaxScenes), "scene table");
    PT_TRY_HIP(d_views_.allocate(sizeof(ViewRecord) * kViewRing), "view ring");
    PT_TRY_HIP(hipMemsetAsync(d_scene_table_.as<void>(), 0, sizeof(SceneDev) * kMaxScenes, stream_), "memset");
    PT_TRY_HIP(hipMemsetAsync(d_views_.as<void>(), 0, sizeof(ViewRecord) * kViewRing, stream_), "memset");
    PT_TRY_HIP(hipMemsetAsync(d_ctr_.as<void>(), 0, sizeof(Counters), stream_), "memset");
    PT_TRY_HIP(hipMemsetAsync(d_overflow_.as<void>(), 0, 64, stream_), "memset");
    PT_TRY_HIP(hipStreamSynchronize(stream_), "sync");
    // the allocation-free trace's scratch on this device (an address lookup, no allocation)
    PT_TRY_HIP(noalloc_scratch_address(reinterpret_cast<void**>(&noalloc_scratch_)), "hipGetSymbolAddress(scratch)");
    return PT_OK;
}

void Renderer::invalidate_graph() {
    for (GraphEntry& g : graphs_) {
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
        if (g.graph) (void)hipGraphDestroy(g.graph);
    }
    graphs_.clear();
}

PathState Renderer::path_state(const LoopMode* mode) const {
    PathState p{};
    if (mode && mode->guided) {
        p.guide = guide_dev();
        p.guide.cfg.guide_fraction = mode->guide_fraction;
        p.guide.cfg.rec_mask = (1u << mode->rec_shift) - 1u;
    }
    for (int k = 0; k < 2; k++) {
        p.ray_o[k] = d_ray_o_[k].as<float4>();
        p.ray_d[k] = d_ray_d_[k].as<float4>();
        p.thr[k] = d_thr_[k].as<float4>();
        p.meta[k] = d_meta_[k].as<uint4>();
    }
    p.hit = d_hit_.as<float4>();
    p.qid = d_qid_.as<uint8_t>();
    for (int i = 0; i < kQueueSlots; i++) p.q[i] = d_q_[i].as<uint32_t>();
    p.sh_o = d_sh_o_.as<float4>();
    p.sh_d = d_sh_d_.as<float4>();
    p.sh_L = d_sh_L_.as<float4>();
    p.ctr = d_ctr_.as<Counters>();
    p.fp = d_fp_.as<FrameParams>();
    p.views = d_views_.as<ViewRecord>();
    p.scenes = d_scene_table_.as<SceneDev>();
    p.accum = d_accum_.as<float4>();
    p.framebuffer = d_frame_.as<uint32_t>();
    p.capacity = pool_capacity_;
    p.pixels = accum_capacity_;
    return p;
}

SceneDev ResidentScene::dev() const {
    SceneDev s{};
    s.nodes = nodes.as<void>();
    s.qnodes = quantized ? qnodes.as<void>() : nullptr;
    s.qnodes8 = depth8 ? qnodes8.as<void>() : nullptr;
    s.tris = tris.as<TriRecord>();
    s.tris8 = depth8 ? tris8.as<TriRecord>() : nullptr;
    s.node8_count = depth8 ? (uint32_t)(qnodes8.bytes() / sizeof(Bvh8QNode)) : 0u;
    s.shade = shade.as<ShadeRecord>();
    s.materials = materials.as<MaterialRecord>();
    s.lights = lights.as<LightRecord>();
    s.bins = bins.as<AliasBin>();
    s.light_count = (uint32_t)host.lights.size();
    s.tri_count = host.tri_count;
    s.node_width = host.bvh.width;
    s.node_quantized = quantized;
    for (int k = 0; k < 2; k++) s.top[k] = top[k].bytes() ? top[k].as<void>() : nullptr;
    if (!quantized) s.top[1] = nullptr;
    return s;
}

// Path slots for a run of `samples` samples: the tuning's fixed size, or, when
// it is 0 (automatic), the next power of two of the largest run so far within
// [kAutoPoolMin, kAutoPoolMax].  A larger pool keeps more paths in flight, so
// fewer bounce iterations carry the end of a frame: the 983k-triangle frame's
// 132.7M samples take 13 iterations at 2^25 slots and 8 at 2^27 (9,000 ->
// 9,810 Mrays/s); 2^27 slots are 28 GB of path state.  The automatic pool only
// grows, so alternating small and large runs does not reallocate, unless
// `exact` asks for the size itself.
pt_status Renderer::ensure_pool(uint64_t samples, bool exact) {
    uint32_t P = tuning_.pool_size;
    if (P == 0) {
        P = kAutoPoolMin;
        while (P < kAutoPoolMax && P < samples) P <<= 1;
        if (!exact && d_hit_.bytes() > 0 && pool_capacity_ >= P) return PT_OK;
    }
    if (P == pool_capacity_ && d_hit_.bytes() > 0) return PT_OK;
    invalidate_graph();
    const size_t f4 = sizeof(float4) * (size_t)P, u1 = sizeof(uint32_t) * (size_t)P;
    DeviceBuffer* f4s[] = {&d_ray_o_[0], &d_ray_o_[1], &d_ray_d_[0], &d_ray_d_[1], &d_thr_[0], &d_thr_[1],
                           &d_meta_[0], &d_meta_[1], &d_hit_, &d_sh_o_, &d_sh_d_, &d_sh_L_};
    for (DeviceBuffer* b : f4s) {
        b->release();
        PT_TRY_HIP(b->allocate(f4), "path state allocation");
    }
    for (int i = 0; i < kQueueSlots; i++) {
        d_q_[i].release();
        if (i < kQueueCount) PT_TRY_HIP(d_q_[i].allocate(u1), "material queue allocation");
    }
    d_qid_.release();
    PT_TRY_HIP(d_qid_.allocate((size_t)P), "queue id allocation");
    pool_capacity_ = P;
    return PT_OK;
}

// Accumulation and tone-mapped pixels for `pixels` pixels over all views of a
// run.  Grows only; the contents are cleared by each run.
pt_status Renderer::ensure_accum(uint32_t pixels) {
    if (pixels <= accum_capacity_ && d_accum_.bytes() > 0) return PT_OK;
    invalidate_graph();
    hdr_valid_ = false;
    const size_t n = std::max<size_t>(pixels, 1);
    d_accum_.release();
    d_frame_.release();
    accum_capacity_ = 0;
    PT_TRY_HIP(d_accum_.allocate(n * sizeof(float4)), "accumulation buffer");
    PT_TRY_HIP(d_frame_.allocate(n * sizeof(uint32_t)), "framebuffer");
    PT_TRY_HIP(hipMemsetAsync(d_accum_.as<void>(), 0, n * sizeof(float4), stream_), "memset");
    PT_TRY_HIP(hipMemsetAsync(d_frame_.as<void>(), 0, n * sizeof(uint32_t), stream_), "memset");
    PT_TRY_HIP(hipStreamSynchronize(stream_), "sync");
    accum_capacity_ = (uint32_t)n;
    return PT_OK;
}

pt_status Renderer::set_frame_size(uint32_t w, uint32_t h) {
    if (frame_locked_here()) return fail(PT_ERROR_INVALID_STATE, "this thread holds the frame lock (pt_frame_lock)");
    if (w == width_ && h == height_ && h_frame_[0]) return PT_OK;
    const size_t n = std::max<size_t>((size_t)w * h, 1);
    std::scoped_lock lock(frame_mutex_);
    if (h_frame_capacity_ < n) {
        for (uint32_t*& b : h_frame_) {
            if (b) (void)hipHostFree(b);
            b = nullptr;
        }
        h_frame_capacity_ = 0;
        width_ = height_ = 0;
        for (uint32_t*& b : h_frame_)
            PT_TRY_HIP(hipHostMalloc((void**)&b, n * sizeof(uint32_t), hipHostMallocDefault), "pinned framebuffer");
        h_frame_capacity_ = n;
    }
    std::memset(h_frame_[h_front_], 0, (size_t)w * h * sizeof(uint32_t));
    width_ = w;
    height_ = h;
    hdr_valid_ = false;
    return PT_OK;
}

// Copies the tone-mapped frame of the last run into the back buffer and makes
// it the front one.  Only the rendering thread writes the back buffer; readers
// hold frame_mutex_ while they copy the front one.
pt_status Renderer::publish_frame() {
    if (frame_locked_here()) return fail(PT_ERROR_INVALID_STATE, "this thread holds the frame lock (pt_frame_lock)");
    const size_t npix = (size_t)width_ * height_;
    uint32_t* back = h_frame_[h_front_ ^ 1];
    PT_TRY_HIP(hipMemcpyAsync(back, d_frame_.as<void>(), npix * sizeof(uint32_t), hipMemcpyDeviceToHost, stream_), "framebuffer");
    PT_TRY_HIP(hipStreamSynchronize(stream_), "frame sync");
    std::scoped_lock lock(frame_mutex_);
    h_front_ ^= 1;
    return PT_OK;
}

namespace {
#ifndef PT_CLOSEST_CHUNK
#define PT_CLOSEST_CHUNK 1024
#endif
#ifndef PT_CLOSEST_CHUNK_MULTI
#define PT_CLOSEST_CHUNK_MULTI 512
#endif
#ifndef PT_ANY_CHUNK
#define PT_ANY_CHUNK 256
#endif
constexpr uint32_t kClosestHitChunk = PT_CLOSEST_CHUNK;
constexpr uint32_t kClosestHitChunkMulti = PT_CLOSEST_CHUNK_MULTI;
constexpr uint32_t kAnyHitChunk = PT_ANY_CHUNK;
}  // namespace

pt_status Renderer::configure_trace() {
    invalidate_graph();
    TraceConfig cfg;
    cfg.block = tuning_.trace_block;
    cfg.lds_stack = cfg.block >= 256 ? 8 : 16;
    cfg.refill = tuning_.trace_refill;
    if (!trace_config_supported(cfg.block, cfg.lds_stack))
        return fail(PT_ERROR_INVALID_ARGUMENT, "unsupported traversal block size " + std::to_string(cfg.block));
    // per traversal and kernel form (one scene or several per launch), the
    // smallest occupancy over the kernels the resident scenes use
    int occ[2][2];   // [multi][any]
    for (int m = 0; m < 2; m++)
        for (int a = 0; a < 2; a++) {
            occ[m][a] = 1 << 20;
            bool any_scene = false;
            for (const auto& sc : scenes_) {
                if (!sc) continue;
                any_scene = true;
                const int kind = (sc->depth8 && (a || (sc->quantized & 8u))) ? 2 : (sc->quantized & (a ? 2u : 1u)) ? 1 : 0;
                occ[m][a] = std::min(occ[m][a], trace_occupancy(cfg, a != 0, false, kind, m != 0));
            }
            if (!any_scene) occ[m][a] = trace_occupancy(cfg, a != 0, false, 0, m != 0);
            occ[m][a] = std::max(1, occ[m][a]);
        }
    const uint32_t cus = (uint32_t)std::max(1, cu_count_);
    uint32_t blocks[2][3];   // [multi][closest, any hit overlapped, any hit alone]
    for (int m = 0; m < 2; m++) {
        int bc = occ[m][0], ba = occ[m][1], alone = occ[m][1];
        if (tuning_.trace_blocks_per_cu) {
            bc = ba = alone = (int)tuning_.trace_blocks_per_cu;
        } else {
            // the overlapped any-hit grid leaves one wave per SIMD (four per CU)
            // free, so that the closest-hit launch it overlaps in the bounce loop
            // starts at once (production frame: 7.8 -> 8.15 Grays/s at 256-thread
            // blocks); alone it takes the whole occupancy
            const int wpb = (int)cfg.block / 64;
            ba = std::max(1, (ba * wpb - 4) / wpb);
        }
        if (tuning_.shadow_blocks_per_cu) ba = alone = (int)tuning_.shadow_blocks_per_cu;
        blocks[m][0] = (uint32_t)bc * cus;
        blocks[m][1] = (uint32_t)ba * cus;
        blocks[m][2] = (uint32_t)alone * cus;
    }
    // worst-case stack over every resident scene: a level of a width-W BVH
    // pushes at most W-1 siblings
    stack_total_ = 3u + 2u;
    for (const auto& sc : scenes_)
        if (sc) {
            stack_total_ = std::max(stack_total_, (sc->host.bvh.width - 1u) * sc->host.bvh.depth + 2u);
            if (sc->depth8) stack_total_ = std::max(stack_total_, 7u * sc->depth8 + 2u);
        }
    cfg.spill_depth = stack_total_ > cfg.lds_stack ? stack_total_ - cfg.lds_stack : 0u;
    // two spill areas, each sized for the larger of its two grids: the any-hit
    // traversal of one bounce runs concurrently with the next bounce's
    // closest-hit traversal
    const size_t spill_closest = (size_t)cfg.spill_depth * std::max(blocks[0][0], blocks[1][0]) * cfg.block;
    const size_t spill_any = (size_t)cfg.spill_depth *
                             std::max({blocks[0][1], blocks[1][1], blocks[0][2], blocks[1][2]}) * cfg.block;
    PT_TRY_HIP(d_spill_.allocate((spill_closest + spill_any) * sizeof(uint2)), "traversal stack spill area");
    cfg.stack_cap = cfg.lds_stack + cfg.spill_depth;
    if (tuning_.stack_limit) cfg.stack_cap = std::min(cfg.stack_cap, tuning_.stack_limit);
    // renders report an overflow in the frame's counters (cleared by the
    // frame's first kernel, read after the frame); the hooks use their own flag
    cfg.overflow = reinterpret_cast<unsigned int*>(d_ctr_.as<char>() + offsetof(Counters, stack_overflow));
    TraceConfig* out[2][3] = {{&trace_, &shadow_trace_, &shadow_alone_},
                              {&trace_multi_, &shadow_trace_multi_, &shadow_alone_multi_}};
    for (int m = 0; m < 2; m++)
        for (int a = 0; a < 3; a++) {
            *out[m][a] = cfg;
            out[m][a]->blocks = blocks[m][a];
            out[m][a]->spill = d_spill_.as<uint2>() + (a ? spill_closest : 0);
            // largest ray chunk a wave claims with one atomic on the launch's
            // work counter, a single address all eight XCDs contend for: over
            // fp32 BVH4 nodes the closest-hit launch (16.7M rays of the
            // production frame) spent a fifth of its time in those claims at
            // 256 rays (12.8 -> 9.9 ms per frame at 512); over the compressed
            // 8-wide tree the production frame takes 27.4 ms at 512 and 26.9 at
            // 1,024, while the several-scene launches of the serving loop run
            // 1 % faster at 512; the any-hit launch's fewer, longer rays
            // balance better with smaller chunks (128 to 512 equal)
            out[m][a]->chunk = a ? kAnyHitChunk : (m ? kClosestHitChunkMulti : kClosestHitChunk);
        }
    hook_trace_ = trace_;
    hook_trace_.overflow = d_overflow_.as<unsigned int>();
    return PT_OK;
}

pt_status Renderer::InitializeRenderingBuffers(uint32_t width, uint32_t height) {
    if (IsRendering()) return fail(PT_ERROR_INVALID_STATE, "Renderer is already rendering");
    width = std::min(std::max(width, kMinWidth), kMaxWidth);
    height = std::min(std::max(height, kMinHeight), kMaxHeight);
    PT_TRY(ensure_pool((uint64_t)width * height));
    PT_TRY(ensure_accum(width * height));
    PT_TRY(set_frame_size(width, height));
    return PT_OK;
}

// Builds the host scene (validation, BVH, light table), uploads its records
// and publishes them in slot `id` of the device scene table.
pt_status Renderer::upload_scene(uint32_t id, const pt_scene_desc& desc) {
    BvhBuildOptions bo;
    bo.max_leaf_size = tuning_.bvh_max_leaf;
    bo.traversal_cost = tuning_.bvh_traversal_cost;
    bo.width = tuning_.bvh_width;
    auto rs = std::make_unique<ResidentScene>();
    SceneError err{PT_OK, ""};
    const bool gpu_bvh = tuning_.bvh_builder == PT_BVH_BUILDER_GPU && desc.triangle_count > 0;
    // shadow rays on 8-wide quantized nodes: a collapse of the host's binary tree
    bo.wide8 = (tuning_.bvh_quantized & 4u) != 0 && bo.width == 4 && !gpu_bvh;
    if (!build_host_scene(desc, bo, rs->host, err, !gpu_bvh)) return fail(err.status, err.message);
    HostScene& h = rs->host;
    auto upload = [&](DeviceBuffer& b, const void* src, size_t bytes, const char* what) -> pt_status {
        PT_TRY_HIP(b.allocate(bytes), what);
        if (bytes) PT_TRY_HIP(hipMemcpyAsync(b.as<void>(), src, bytes, hipMemcpyHostToDevice, stream_), what);
        return PT_OK;
    };
    if (gpu_bvh) {
        // positions and material ids to the device, then the linear-BVH build there
        const uint32_t T = desc.triangle_count;
        DeviceBuffer pos, mids, nodes;
        PT_TRY(upload(pos, desc.positions, (size_t)T * 9 * sizeof(float), "positions"));
        PT_TRY(upload(mids, desc.material_ids, (size_t)T * sizeof(uint32_t), "material ids"));
        PT_TRY(upload(rs->materials, h.materials.data(), h.materials.size() * sizeof(MaterialRecord), "materials"));
        const uint32_t cap = gpu_bvh4_node_bound(T);
        PT_TRY_HIP(nodes.allocate((size_t)cap * sizeof(Bvh4Node)), "BVH nodes");
        PT_TRY_HIP(rs->tris.allocate((size_t)T * sizeof(TriRecord)), "triangle records");
        PT_TRY_HIP(rs->order.allocate((size_t)T * sizeof(uint32_t)), "BVH order");
        // a Morton tree has no SAH to size its leaves: leaves of at most two
        // triangles traced fastest (production frame: 6479 Mrays/s against
        // 6084 with four and 5314 with eight; the host SAH tree 6701)
        const uint32_t leaf = std::min(bo.max_leaf_size, kGpuBvhLeaf);
        GpuBvhStats st;
        PT_TRY_HIP(gpu_build_bvh4(stream_, pos.as<float>(), mids.as<uint32_t>(), rs->materials.as<MaterialRecord>(), T,
                                  leaf, nodes.as<Bvh4Node>(), cap, rs->tris.as<TriRecord>(),
                                  rs->order.as<uint32_t>(), st),
                   "GPU BVH build");
        // keep only the nodes written
        PT_TRY_HIP(rs->nodes.allocate((size_t)st.node_count * sizeof(Bvh4Node)), "BVH nodes");
        PT_TRY_HIP(hipMemcpyAsync(rs->nodes.as<void>(), nodes.as<void>(), (size_t)st.node_count * sizeof(Bvh4Node),
                                  hipMemcpyDeviceToDevice, stream_),
                   "BVH nodes");
        PT_TRY_HIP(hipStreamSynchronize(stream_), "BVH nodes");
        h.bvh.width = 4;
        h.bvh.depth = st.depth;
        h.bvh.leaf_count = st.leaf_count;
        h.bvh.max_leaf = st.max_leaf;
        h.bvh.device_nodes = st.node_count;
    } else if (h.bvh.width == 2) {
        const std::vector<Bvh2Node> n2 = to_bvh2(h.bvh);
        PT_TRY(upload(rs->nodes, n2.data(), n2.size() * sizeof(Bvh2Node), "BVH nodes"));
        PT_TRY_HIP(hipStreamSynchronize(stream_), "BVH nodes");
    } else {
        PT_TRY(upload(rs->nodes, h.bvh.nodes.data(), h.bvh.nodes.size() * sizeof(Bvh4Node), "BVH nodes"));
    }
    if (tuning_.bvh_quantized && h.bvh.width == 4) {
        // 64-byte quantized copy of the 4-wide nodes for traversal
        if (gpu_bvh) {
            const uint32_t nn = h.bvh.device_nodes;
            PT_TRY_HIP(rs->qnodes.allocate((size_t)nn * sizeof(Bvh4QNode)), "quantized BVH nodes");
            PT_TRY_HIP(gpu_quantize_bvh4(stream_, rs->nodes.as<Bvh4Node>(), nn, rs->qnodes.as<Bvh4QNode>()),
                       "BVH node quantization");
        } else {
            std::vector<Bvh4QNode> q(h.bvh.nodes.size());
            for (size_t i = 0; i < q.size(); i++) quantize_bvh4_node(h.bvh.nodes[i], q[i]);
            PT_TRY(upload(rs->qnodes, q.data(), q.size() * sizeof(Bvh4QNode), "quantized BVH nodes"));
            PT_TRY_HIP(hipStreamSynchronize(stream_), "quantized BVH nodes");
        }
        rs->quantized = tuning_.bvh_quantized;
    }
    if (!h.bvh.qnodes8.empty()) {
        // the 8-wide quantized collapse of the same binary tree, for shadow rays
        PT_TRY(upload(rs->qnodes8, h.bvh.qnodes8.data(), h.bvh.qnodes8.size() * sizeof(Bvh8QNode),
                      "8-wide quantized BVH nodes"));
        PT_TRY(upload(rs->tris8, h.tris8.data(), h.tris8.size() * sizeof(TriRecord), "8-wide tree triangle records"));
        PT_TRY_HIP(hipStreamSynchronize(stream_), "8-wide quantized BVH nodes");
        rs->depth8 = h.bvh.depth8;
    }
    // top-of-tree tables, which the traversal blocks keep in LDS
    const uint64_t node_count = gpu_bvh ? h.bvh.device_nodes : h.bvh.nodes.size();
    if (node_count > kMaxNodeIndex) return fail(PT_ERROR_OUT_OF_RANGE, "the BVH has more than 2^30 nodes");
    auto build_top = [&](DeviceBuffer& table, int kind, const DeviceBuffer& nodes) -> hipError_t {
        hipError_t e = table.allocate(top_table_bytes(kind));
        return e != hipSuccess ? e : launch_build_top(stream_, kind, nodes.as<void>(), table.as<void>());
    };
    PT_TRY_HIP(build_top(rs->top[0], h.bvh.width == 2 ? kNodeBvh2 : kNodeBvh4, rs->nodes), "top-of-tree table");
    if (rs->quantized) PT_TRY_HIP(build_top(rs->top[1], kNodeBvh4Q, rs->qnodes), "top-of-tree table");
    if (!gpu_bvh) {
        PT_TRY(upload(rs->tris, h.tris.data(), h.tris.size() * sizeof(TriRecord), "triangle records"));
        PT_TRY(upload(rs->materials, h.materials.data(), h.materials.size() * sizeof(MaterialRecord), "materials"));
    }
    PT_TRY(upload(rs->shade, h.shade.data(), h.shade.size() * sizeof(ShadeRecord), "shading records"));
    PT_TRY(upload(rs->lights, h.lights.data(), h.lights.size() * sizeof(LightRecord), "lights"));
    PT_TRY(upload(rs->bins, h.bins.data(), h.bins.size() * sizeof(AliasBin), "alias table"));
    const SceneDev d = rs->dev();
    PT_TRY_HIP(launch_write_words(stream_, d_scene_table_.as<SceneDev>() + id, &d, sizeof(SceneDev)), "scene table");
    PT_TRY_HIP(hipStreamSynchronize(stream_), "scene upload");
    invalidate_graph();
    scenes_[id] = std::move(rs);
    return configure_trace();
}

pt_status Renderer::InitializeSceneBuffers(const pt_scene_desc& desc) {
    if (IsRendering()) return fail(PT_ERROR_INVALID_STATE, "Renderer is already rendering");
    return upload_scene(0, desc);
}

pt_status Renderer::AddScene(const pt_scene_desc& desc, uint32_t& id) {
    if (IsRendering()) return fail(PT_ERROR_INVALID_STATE, "Renderer is already rendering");
    for (uint32_t k = 1; k < kMaxScenes; k++) {
        if (scenes_[k]) continue;
        PT_TRY(upload_scene(k, desc));
        id = k;
        return PT_OK;
    }
    return fail(PT_ERROR_OUT_OF_RANGE, "all " + std::to_string(kMaxScenes - 1) + " scene slots are in use");
}

pt_status Renderer::RemoveScene(uint32_t id) {
    if (IsRendering()) return fail(PT_ERROR_INVALID_STATE, "Renderer is already rendering");
    if (id >= kMaxScenes || !scenes_[id]) return fail(PT_ERROR_INVALID_ARGUMENT, "no resident scene " + std::to_string(id));
    invalidate_graph();
    const SceneDev zero{};
    PT_TRY_HIP(launch_write_words(stream_, d_scene_table_.as<SceneDev>() + id, &zero, sizeof(SceneDev)), "scene table");
    PT_TRY_HIP(hipStreamSynchronize(stream_), "scene table");
    scenes_[id].reset();
    return configure_trace();
}

pt_status Renderer::Clear(uint8_t r, uint8_t g, uint8_t b, uint8_t a) {
    if (IsRendering()) return fail(PT_ERROR_INVALID_STATE, "Cannot clear renderer while rendering");
    if (frame_locked_here()) return fail(PT_ERROR_INVALID_STATE, "this thread holds the frame lock (pt_frame_lock)");
    const uint32_t c = (uint32_t)r | ((uint32_t)g << 8) | ((uint32_t)b << 16) | ((uint32_t)a << 24);
    std::scoped_lock lock(frame_mutex_);
    if (h_frame_[h_front_]) std::fill(h_frame_[h_front_], h_frame_[h_front_] + (size_t)width_ * height_, c);
    return PT_OK;
}

pt_status Renderer::join_shadow() {
    if (shadow_pending_ < 0) return PT_OK;
    PT_TRY_HIP(hipStreamWaitEvent(stream_, join_events_[shadow_pending_], 0), "join");
    shadow_pending_ = -1;
    return PT_OK;
}

pt_status Renderer::launch_iteration(int cur, bool count, const LoopMode& mode, bool fork) {
    const PathState p = path_state(&mode);
    const SceneDev sc = scene_dev(mode.scene);
    const uint32_t grid = (uint32_t)std::max(1, cu_count_) * 8u;
    const TraceConfig& tc = mode.trace_multi ? trace_multi_ : trace_;
    const TraceConfig& ta = fork ? (mode.trace_multi ? shadow_trace_multi_ : shadow_trace_)
                                 : (mode.trace_multi ? shadow_alone_multi_ : shadow_alone_);
    PT_TRY_HIP(launch_closest(stream_, p, sc, cur, count, mode.trace_multi, tc), "closest-hit launch");
    PT_TRY_HIP(launch_bin(stream_, p, cur, bin_grid()), "binning launch");
    // the shadow queue and the guide records are rewritten by this shade
    PT_TRY(join_shadow());
    PT_TRY_HIP(launch_shade(stream_, p, sc, cur, mode.shade_multi, mode.guided, grid, 256), "shade launch");
    if (fork) {
        PT_TRY_HIP(hipEventRecord(fork_events_[cur], stream_), "fork");
        PT_TRY_HIP(hipStreamWaitEvent(side_stream_, fork_events_[cur], 0), "fork");
        PT_TRY_HIP(launch_shadow(side_stream_, p, sc, cur, count, mode.trace_multi, ta), "shadow launch");
        PT_TRY_HIP(hipEventRecord(join_events_[cur], side_stream_), "join");
        shadow_pending_ = cur;
    } else {
        PT_TRY_HIP(launch_shadow(stream_, p, sc, cur, count, mode.trace_multi, ta), "shadow launch");
    }
    PT_TRY_HIP(launch_generate(stream_, p, false, cur ^ 1, grid, 256), "regenerate launch");
    PT_TRY_HIP(launch_prepare(stream_, p, cur), "prepare launch");
    return PT_OK;
}

pt_status Renderer::ensure_graph(const LoopMode& mode, hipGraphExec_t& exec) {
    const int key = mode.key();
    for (const GraphEntry& g : graphs_)
        if (g.key == key) {
            exec = g.exec;
            return PT_OK;
        }
    PT_TRY_HIP(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal), "graph capture");
    // kGraphIterations bounces, each any-hit launch overlapping the next
    // bounce's first half; the last one joins before the graph ends
    shadow_pending_ = -1;
    pt_status s = PT_OK;
    for (int i = 0; i < kGraphIterations && s == PT_OK; i++) s = launch_iteration(i & 1, false, mode, true);
    if (s == PT_OK) s = join_shadow();
    shadow_pending_ = -1;
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(stream_, &g);
    if (s != PT_OK) {
        if (g) (void)hipGraphDestroy(g);
        return s;
    }
    PT_TRY_HIP(e, "graph capture");
    hipGraphExec_t ge = nullptr;
    e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        (void)hipGraphDestroy(g);
        return check_hip(e, "graph instantiate");
    }
    graphs_.push_back(GraphEntry{key, g, ge});
    exec = ge;
    return PT_OK;
}

namespace {
struct ClampedSettings {
    uint32_t width, height, spp, axis, depth, rr;
};
ClampedSettings clamp_settings(const pt_render_settings& s) {
    ClampedSettings c;
    c.width = std::min(std::max(s.width, kMinWidth), kMaxWidth);
    c.height = std::min(std::max(s.height, kMinHeight), kMaxHeight);
    const uint32_t spp = std::min(std::max(s.samples_per_pixel, kMinSpp), kMaxSpp);
    c.depth = std::min(std::max(s.path_depth, kMinDepth), kMaxDepth);
    c.axis = (uint32_t)std::sqrt((double)spp);   // Renderer.cpp:366-367
    while ((uint64_t)(c.axis + 1) * (c.axis + 1) <= spp) c.axis++;
    while ((uint64_t)c.axis * c.axis > spp) c.axis--;
    c.spp = c.axis * c.axis;
    c.rr = s.rr_start_depth ? s.rr_start_depth : kDefaultRrStart;
    return c;
}
}  // namespace

void frame_size(const pt_render_settings& s, uint32_t& width, uint32_t& height) {
    const ClampedSettings c = clamp_settings(s);
    width = c.width;
    height = c.height;
}

// View record of `v` with settings clamped as in Renderer.cpp:354-372; the
// caller fills sample_begin and accum_offset.
pt_status Renderer::make_view(const pt_view& v, ViewRecord& rec, uint32_t& depth) {
    if (v.scene >= kMaxScenes || !scenes_[v.scene])
        return fail(PT_ERROR_INVALID_ARGUMENT, "view names scene " + std::to_string(v.scene) + ", which is not resident");
    const ClampedSettings c = clamp_settings(v.settings);
    std::memset(&rec, 0, sizeof(rec));
    rec.cam = make_camera_basis(v.camera, (float)c.width / (float)c.height, v.state);
    rec.width = c.width;
    rec.height = c.height;
    rec.spp = c.spp;
    rec.spp_axis = c.axis;
    rec.depth_limit = c.depth;
    rec.use_nee = v.settings.use_nee ? 1u : 0u;
    rec.rr_start = c.rr;
    rec.scene = v.scene;
    rec.seed = v.settings.seed;
    rec.sample_count = (uint64_t)c.spp * c.width * c.height;
    rec.inv_spp = 1.0f / (float)c.spp;
    depth = c.depth;
    return PT_OK;
}

// Renders `views` (sample_begin and accum_offset consecutive from 0) in one
// wavefront: every kernel of the bounce loop draws paths of all views from
// the one pool.  Leaves the tone-mapped pixels in d_frame_ and the radiance
// sums in d_accum_.
pt_status Renderer::run_views(const std::vector<ViewRecord>& views, uint32_t max_depth, uint32_t flags,
                              pt_frame_stats* stats, const std::atomic<bool>* stop, const RunOptions& opt) {
    const uint32_t n = (uint32_t)views.size();
    if (n == 0 || n > kViewRing) return fail(PT_ERROR_INVALID_ARGUMENT, "a batch holds 1 to " + std::to_string(kViewRing) + " views");
    const ViewRecord& last = views.back();
    const uint64_t total_samples = last.sample_begin + last.sample_count;
    const uint64_t total_pixels64 = (uint64_t)last.accum_offset + (uint64_t)last.width * last.height;
    if (total_pixels64 > 0x7FFFFFFFull) return fail(PT_ERROR_OUT_OF_RANGE, "the batch holds more than 2^31 pixels");
    const uint32_t total_pixels = (uint32_t)total_pixels64;
    PT_TRY(ensure_pool(total_samples));
    PT_TRY(ensure_accum(total_pixels));

    LoopMode mode;
    mode.scene = views[0].scene;
    mode.shade_multi = n > 1;
    for (const ViewRecord& v : views) {
        mode.trace_multi = mode.trace_multi || v.scene != mode.scene;
        if (scenes_[v.scene]->host.bvh.width != scenes_[mode.scene]->host.bvh.width ||
            scenes_[v.scene]->quantized != scenes_[mode.scene]->quantized ||
            (scenes_[v.scene]->depth8 != 0) != (scenes_[mode.scene]->depth8 != 0))
            return fail(PT_ERROR_INVALID_ARGUMENT, "the scenes of a batch must share one BVH node format");
    }
    mode.guided = opt.guided;
    mode.guide_fraction = opt.guided ? opt.guide_fraction : 0.0f;
    mode.rec_shift = opt.guided ? std::min(opt.rec_shift, 31u) : 0u;
    if (mode.guided && (mode.trace_multi || mode.shade_multi))
        return fail(PT_ERROR_INVALID_ARGUMENT, "guided rendering takes one view");

    FrameParams& fp = *h_fp_;
    std::memset(&fp, 0, sizeof(fp));
    fp.total_samples = total_samples;
    fp.view_lo = 0;
    fp.view_hi = n;
    for (uint32_t b = 0; b <= kMaxDepthLimit; b++) fp.bounce_jump[b] = pcg_jump(bounce_draw_base(b));

    const bool profile = (flags & PT_FRAME_PROFILE) != 0;
    const bool count = (flags & PT_FRAME_INSTRUMENT) != 0;
    const bool use_graph = tuning_.use_graph && !profile && !count && !(flags & PT_FRAME_NO_GRAPH);
    hipGraphExec_t graph_exec = nullptr;
    if (use_graph) PT_TRY(ensure_graph(mode, graph_exec));

    hipEvent_t t_begin = nullptr, t_end = nullptr;
    PT_TRY_HIP(hipEventCreate(&t_begin), "event");
    PT_TRY_HIP(hipEventCreate(&t_end), "event");
    struct EventGuard {
        hipEvent_t a, b;
        ~EventGuard() {
            if (a) (void)hipEventDestroy(a);
            if (b) (void)hipEventDestroy(b);
        }
    } guard{t_begin, t_end};

    const PathState p = path_state(&mode);
    const SceneDev sc = scene_dev(mode.scene);
    const uint32_t grid = (uint32_t)std::max(1, cu_count_) * 8u;
    PT_TRY_HIP(launch_write_words(stream_, d_views_.as<void>(), views.data(), n * sizeof(ViewRecord)), "view records");
    PT_TRY_HIP(launch_write_words(stream_, d_fp_.as<void>(), h_fp_, sizeof(FrameParams)), "frame params");
    PT_TRY_HIP(hipEventRecord(t_begin, stream_), "event");
    PT_TRY_HIP(launch_frame_begin(stream_, p), "frame begin");
    if (opt.clear) PT_TRY_HIP(launch_fill_accum(stream_, p, 0, total_pixels), "clear accumulation");
    PT_TRY_HIP(launch_generate(stream_, p, true, 0, grid, 256), "initialize paths");

    // Every path ends within `depth` iterations and slots are refilled at once,
    // so the loop needs at most (ceil(total/P)+1)*depth iterations.
    const uint64_t waves = (total_samples + pool_capacity_ - 1) / pool_capacity_ + 1;
    const uint64_t max_iters = 2 * waves * (uint64_t)(max_depth + 1) + 16;

    double closest_ms = 0, shadow_ms = 0, shade_ms = 0, regen_ms = 0, bin_ms = 0;
    uint32_t closest_launches = 0;
    if (profile) {
        hipEvent_t ev[7];
        for (auto& e : ev) PT_TRY_HIP(hipEventCreate(&e), "event");
        struct Guard7 {
            hipEvent_t* e;
            ~Guard7() {
                for (int i = 0; i < 7; i++) (void)hipEventDestroy(e[i]);
            }
        } g7{ev};
        int cur = 0;
        queue_history_.clear();
        for (uint64_t it = 0;; it++) {
            if (it > max_iters) return fail(PT_ERROR_DEVICE, "wavefront did not terminate");
            PT_TRY_HIP(hipEventRecord(ev[0], stream_), "event");
            PT_TRY_HIP(launch_closest(stream_, p, sc, cur, false, mode.trace_multi,
                                      mode.trace_multi ? trace_multi_ : trace_), "closest-hit launch");
            PT_TRY_HIP(hipEventRecord(ev[1], stream_), "event");
            PT_TRY_HIP(launch_bin(stream_, p, cur, bin_grid()), "binning launch");
            PT_TRY_HIP(hipEventRecord(ev[2], stream_), "event");
            PT_TRY_HIP(launch_shade(stream_, p, sc, cur, mode.shade_multi, mode.guided, grid, 256), "shade launch");
            PT_TRY_HIP(hipEventRecord(ev[3], stream_), "event");
            PT_TRY_HIP(launch_shadow(stream_, p, sc, cur, false, mode.trace_multi,
                                     mode.trace_multi ? shadow_alone_multi_ : shadow_alone_), "shadow launch");
            PT_TRY_HIP(hipEventRecord(ev[4], stream_), "event");
            PT_TRY_HIP(launch_generate(stream_, p, false, cur ^ 1, grid, 256), "regenerate launch");
            PT_TRY_HIP(hipEventRecord(ev[5], stream_), "event");
            // the iteration's queue sizes, before the prepare kernel resets them
            PT_TRY_HIP(hipMemcpyAsync(h_ctr_ring_ + 1, p.ctr, sizeof(Counters), hipMemcpyDeviceToHost, stream_), "counters");
            PT_TRY_HIP(launch_prepare(stream_, p, cur), "prepare launch");
            PT_TRY_HIP(hipMemcpyAsync(h_ctr_ring_, p.ctr, sizeof(Counters), hipMemcpyDeviceToHost, stream_), "counters");
            PT_TRY_HIP(hipEventRecord(ev[6], stream_), "event");
            PT_TRY_HIP(hipEventSynchronize(ev[6]), "sync");
            float ms[5];
            for (int k = 0; k < 5; k++) PT_TRY_HIP(hipEventElapsedTime(&ms[k], ev[k], ev[k + 1]), "elapsed");
            closest_ms += ms[0];
            bin_ms += ms[1];
            shade_ms += ms[2];
            shadow_ms += ms[3];
            regen_ms += ms[4];
            closest_launches++;
            {
                const Counters& q = h_ctr_ring_[1];
                pt_queue_occupancy o{};
                o.extension = q.ray[cur];
                for (int k = 0; k < 4; k++) o.material[k] = q.queue[k];
                o.missed = q.queue[4];
                o.shadow = q.shadow[cur];
                o.next = h_ctr_ring_[0].ray[cur ^ 1];
                o.capacity = pool_capacity_;
                if (o.extension) queue_history_.push_back(o);
            }
            done_samples_.store(h_ctr_ring_[0].launched, std::memory_order_relaxed);
            cur ^= 1;
            if (h_ctr_ring_[0].active == 0) break;
            if (stop && stop->load()) return PT_OK;
        }
    } else {
        // the host reads the counters of the previous graph launch while the
        // current one runs: enough to keep the queue ahead of the device
        // (two launches of lag measured no faster, and overshot the frame's end)
        const int lag = 1;
        int cur = 0;
        for (uint64_t k = 0;; k++) {
            if (k * (use_graph ? kGraphIterations : 1) > max_iters)
                return fail(PT_ERROR_DEVICE, "wavefront did not terminate");
            if (use_graph) {
                PT_TRY_HIP(hipGraphLaunch(graph_exec, stream_), "graph launch");
            } else {
                PT_TRY(launch_iteration(cur, count, mode));
                cur ^= 1;
            }
            const int slot = (int)(k % kRing);
            PT_TRY_HIP(hipMemcpyAsync(h_ctr_ring_ + slot, p.ctr, sizeof(Counters), hipMemcpyDeviceToHost, stream_), "counters");
            PT_TRY_HIP(hipEventRecord(ring_events_[slot], stream_), "event");
            if (k >= (uint64_t)lag) {
                const int old = (int)((k - lag) % kRing);
                PT_TRY_HIP(hipEventSynchronize(ring_events_[old]), "sync");
                done_samples_.store(h_ctr_ring_[old].launched, std::memory_order_relaxed);
                if (h_ctr_ring_[old].active == 0) break;
            }
            if (stop && stop->load()) {
                PT_TRY_HIP(hipStreamSynchronize(stream_), "sync");
                return PT_OK;
            }
        }
    }
    PT_TRY_HIP(launch_postprocess(stream_, p, 0, n, 0, total_pixels), "postprocess");
    PT_TRY_HIP(hipEventRecord(t_end, stream_), "event");
    PT_TRY_HIP(hipMemcpyAsync(h_ctr_ring_, p.ctr, sizeof(Counters), hipMemcpyDeviceToHost, stream_), "counters");
    PT_TRY_HIP(hipStreamSynchronize(stream_), "frame sync");
    const Counters& ctr = h_ctr_ring_[0];
    if (ctr.stack_overflow) return fail(PT_ERROR_DEVICE, "traversal stack overflow");
    float total_ms = 0;
    PT_TRY_HIP(hipEventElapsedTime(&total_ms, t_begin, t_end), "elapsed");
    done_samples_.store(ctr.launched, std::memory_order_relaxed);
    if (stats) {
        std::memset(stats, 0, sizeof(*stats));
        stats->render_ms = total_ms;
        stats->extension_rays = ctr.ext_rays;
        stats->shadow_rays = ctr.shadow_rays;
        stats->samples = ctr.launched;
        stats->iterations = ctr.iterations;
        stats->closest_launches = profile ? closest_launches : ctr.iterations;
        stats->closest_ms = closest_ms;
        stats->shadow_ms = shadow_ms;
        stats->shade_ms = shade_ms;
        stats->regen_ms = regen_ms;
        stats->bin_ms = bin_ms;
        stats->closest_nodes = ctr.closest_nodes;
        stats->closest_tris = ctr.closest_tris;
        stats->shadow_nodes = ctr.shadow_nodes;
        stats->shadow_tris = ctr.shadow_tris;
    }
    if (ctr.invalid_pixels) return fail(PT_ERROR_DEVICE, std::to_string(ctr.invalid_pixels) + " pixels hold non-finite or negative radiance");
    return PT_OK;
}

pt_status Renderer::render_frame(const pt_render_settings& s, const pt_camera& camera, const pt_motion_state& state,
                                 uint32_t flags, pt_frame_stats* stats, const std::atomic<bool>* stop) {
    if (!has_scene()) return fail(PT_ERROR_INVALID_STATE, "scene buffers are not initialized");
    pt_view v{};
    v.scene = 0;
    v.settings = s;
    v.camera = camera;
    v.state = state;
    std::vector<ViewRecord> views(1);
    uint32_t depth = 0;
    PT_TRY(make_view(v, views[0], depth));
    PT_TRY(set_frame_size(views[0].width, views[0].height));
    last_spp_ = views[0].spp;
    hdr_valid_ = false;
    PT_TRY(run_views(views, depth, flags, stats, stop));
    if (stop && stop->load()) return PT_OK;
    PT_TRY(publish_frame());
    hdr_valid_ = true;
    return PT_OK;
}

pt_status Renderer::RenderBatch(const pt_view* views, uint32_t n, float* const* hdr, uint32_t* const* rgba,
                                uint32_t flags, pt_frame_stats* stats) {
    if (IsRendering()) return fail(PT_ERROR_INVALID_STATE, "Renderer is already rendering");
    if (!views || n == 0) return fail(PT_ERROR_INVALID_ARGUMENT, "a batch needs at least one view");
    if (n > kViewRing) return fail(PT_ERROR_INVALID_ARGUMENT, "a batch holds at most " + std::to_string(kViewRing) + " views");
    std::vector<ViewRecord> recs(n);
    uint64_t samples = 0, pixels = 0;
    uint32_t max_depth = 1;
    for (uint32_t i = 0; i < n; i++) {
        uint32_t depth = 0;
        PT_TRY(make_view(views[i], recs[i], depth));
        recs[i].sample_begin = samples;
        recs[i].accum_offset = (uint32_t)std::min<uint64_t>(pixels, 0xFFFFFFFFull);
        samples += recs[i].sample_count;
        pixels += (uint64_t)recs[i].width * recs[i].height;
        max_depth = std::max(max_depth, depth);
    }
    if (pixels > 0x7FFFFFFFull) return fail(PT_ERROR_OUT_OF_RANGE, "the batch holds more than 2^31 pixels");
    hdr_valid_ = false;
    PT_TRY(run_views(recs, max_depth, flags, stats, nullptr));
    for (uint32_t i = 0; i < n; i++) {
        const size_t np = (size_t)recs[i].width * recs[i].height;
        if (rgba && rgba[i])
            PT_TRY_HIP(hipMemcpy(rgba[i], d_frame_.as<uint32_t>() + recs[i].accum_offset, np * sizeof(uint32_t),
                                 hipMemcpyDeviceToHost), "batch framebuffer");
        if (hdr && hdr[i]) {
            PT_TRY_HIP(hipMemcpy(hdr[i], d_accum_.as<float4>() + recs[i].accum_offset, np * sizeof(float4),
                                 hipMemcpyDeviceToHost), "batch HDR");
            const float inv = recs[i].inv_spp;
            for (size_t k = 0; k < 4 * np; k++) hdr[i][k] *= inv;
        }
    }
    return PT_OK;
}

pt_status Renderer::RenderFrameSync(const pt_render_settings& s, const pt_camera& camera, const pt_motion_state& state,
                                    uint32_t flags, pt_frame_stats* stats) {
    if (IsRendering()) return fail(PT_ERROR_INVALID_STATE, "Renderer is already rendering");
    return render_frame(s, camera, state, flags, stats, nullptr);
}

// ---------------------------------------------------------------- async rendering

pt_status Renderer::StartSimulation(const pt_render_settings& s, const pt_camera& camera, const pt_motion_state* states,
                                    uint32_t n, const char* output_folder) {
    if (is_rendering_.load()) return fail(PT_ERROR_INVALID_STATE, "Renderer is already rendering");
    if (!has_scene()) return fail(PT_ERROR_INVALID_STATE, "scene buffers are not initialized");
    if (n > 0 && !states) return fail(PT_ERROR_INVALID_ARGUMENT, "motion states are NULL");
    if (render_thread_.joinable()) render_thread_.join();
    {
        std::scoped_lock lock(error_mutex_);
        parked_error_.reset();
    }
    const ClampedSettings c = clamp_settings(s);
    PT_TRY(ensure_pool((uint64_t)c.spp * c.width * c.height));
    PT_TRY(ensure_accum(c.width * c.height));
    PT_TRY(set_frame_size(c.width, c.height));
    settings_ = s;
    camera_ = camera;
    motion_.assign(states, states + n);
    output_folder_ = output_folder ? output_folder : "";
    total_frames_ = n;
    total_samples_ = (uint64_t)c.spp * c.width * c.height;
    done_frames_.store(0);
    done_samples_.store(0);
    stop_requested_.store(false);
    is_rendering_.store(true);
    try {
        render_thread_ = std::thread([this]() { simulation_loop(0); });
    } catch (...) {
        is_rendering_.store(false);
        return fail(PT_ERROR_INVALID_STATE, "Failed to start render thread");
    }
    return PT_OK;
}

pt_status Renderer::ResumeSimulation(uint32_t start_frame) {
    if (is_rendering_.load()) return fail(PT_ERROR_INVALID_STATE, "Renderer is already rendering");
    if (start_frame >= motion_.size())
        return fail(PT_ERROR_OUT_OF_RANGE, "Start frame is out of range - nothing to render");
    if (render_thread_.joinable()) render_thread_.join();
    done_frames_.store(start_frame);
    done_samples_.store(0);
    stop_requested_.store(false);
    is_rendering_.store(true);
    try {
        render_thread_ = std::thread([this, start_frame]() { simulation_loop(start_frame); });
    } catch (...) {
        is_rendering_.store(false);
        return fail(PT_ERROR_INVALID_STATE, "Failed to start render thread");
    }
    return PT_OK;
}

pt_status Renderer::StopRendering() {
    if (server_) return fail(PT_ERROR_INVALID_STATE, "the server is running: stop it with pt_server_stop");
    // the render thread may be waiting for the frame lock to publish a frame
    if (frame_locked_here()) return fail(PT_ERROR_INVALID_STATE, "this thread holds the frame lock (pt_frame_lock)");
    if (render_thread_.joinable()) {
        stop_requested_.store(true);
        render_thread_.join();
    }
    if (stream_) {
        hipError_t e = hipStreamSynchronize(stream_);
        if (e != hipSuccess) {
            is_rendering_.store(false);
            return check_hip(e, "StopRendering");
        }
    }
    is_rendering_.store(false);
    return PT_OK;
}

void Renderer::simulation_loop(uint32_t start_frame) {
    for (uint32_t f = start_frame; f < total_frames_; f++) {
        pt_status s = render_frame(settings_, camera_, motion_[f], 0, nullptr, &stop_requested_);
        if (s == PT_OK && !stop_requested_.load() && !output_folder_.empty()) {
            const int digits = (int)std::max<size_t>(4, std::to_string(total_frames_ ? total_frames_ - 1 : 0).size());
            char name[64];
            std::snprintf(name, sizeof(name), "frame_%0*u.png", digits, done_frames_.load());
            s = save_frame_png((std::filesystem::path(output_folder_) / name).string());
        }
        if (s != PT_OK) {
            (void)hipStreamSynchronize(stream_);
            std::scoped_lock lock(error_mutex_);
            parked_error_ = last_error_;
            done_samples_.store(0);
            break;
        }
        if (stop_requested_.load()) break;
        done_frames_.fetch_add(1);
    }
    is_rendering_.store(false);
}

void Renderer::GetProgress(pt_progress& p) const {
    std::memset(&p, 0, sizeof(p));
    p.is_rendering = is_rendering_.load() ? 1u : 0u;
    p.done_frames = done_frames_.load();
    p.total_frames = total_frames_;
    p.width = width_;
    p.height = height_;
    p.samples_per_pixel = last_spp_;
    p.samples_per_pixel_axis = (uint32_t)std::lround(std::sqrt((double)last_spp_));
    p.path_depth = std::min(std::max(settings_.path_depth, kMinDepth), kMaxDepth);
    p.done_samples = done_samples_.load();
    p.total_samples = total_samples_;
}

pt_status Renderer::CopyFrame(uint32_t* rgba, size_t n) {
    if (frame_locked_here()) return fail(PT_ERROR_INVALID_STATE, "this thread holds the frame lock (pt_frame_lock)");
    std::scoped_lock lock(frame_mutex_);
    const size_t npix = (size_t)width_ * height_;
    if (!rgba || n < npix) return fail(PT_ERROR_INVALID_ARGUMENT, "frame buffer too small");
    if (npix) std::memcpy(rgba, h_frame_[h_front_], npix * sizeof(uint32_t));
    return PT_OK;
}

// FrameView::Lock() (Renderer.hpp:45-104) without a copy: the front buffer
// under the frame lock until UnlockFrame on the same thread.
pt_status Renderer::LockFrame(const uint32_t** rgba, size_t* n) {
    if (!rgba || !n) return fail(PT_ERROR_INVALID_ARGUMENT, "NULL argument");
    if (frame_locked_here()) return fail(PT_ERROR_INVALID_STATE, "this thread already holds the frame lock");
    frame_mutex_.lock();
    frame_lock_owner_.store(std::this_thread::get_id());
    *rgba = h_frame_[h_front_];
    *n = h_frame_[h_front_] ? (size_t)width_ * height_ : 0;
    return PT_OK;
}

pt_status Renderer::UnlockFrame() {
    if (!frame_locked_here()) return fail(PT_ERROR_INVALID_STATE, "this thread does not hold the frame lock");
    frame_lock_owner_.store(std::thread::id());
    frame_mutex_.unlock();
    return PT_OK;
}

pt_status Renderer::CopyFrameHdr(float* rgba, size_t n) {
    if (IsRendering()) return fail(PT_ERROR_INVALID_STATE, "HDR readback needs a finished render");
    if (!hdr_valid_) return fail(PT_ERROR_INVALID_STATE, "no finished single frame to read back");
    const size_t npix = (size_t)width_ * height_;
    if (!rgba || n < npix) return fail(PT_ERROR_INVALID_ARGUMENT, "HDR buffer too small");
    PT_TRY_HIP(hipMemcpyAsync(rgba, d_accum_.as<void>(), npix * sizeof(float4), hipMemcpyDeviceToHost, stream_), "HDR readback");
    PT_TRY_HIP(hipStreamSynchronize(stream_), "HDR readback");
    const float inv = 1.0f / (float)last_spp_;
    for (size_t i = 0; i < 4 * npix; i++) rgba[i] *= inv;
    return PT_OK;
}

static size_t copy_message(const std::string& m, char* buf, size_t cap) {
    if (buf && cap) {
        const size_t k = std::min(cap - 1, m.size());
        std::memcpy(buf, m.data(), k);
        buf[k] = 0;
    }
    return m.size();
}

size_t Renderer::PeekLastError(char* buf, size_t cap) const {
    std::scoped_lock lock(error_mutex_);
    if (!parked_error_) return 0;
    return copy_message(*parked_error_, buf, cap);
}

size_t Renderer::ConsumeLastError(char* buf, size_t cap) {
    std::scoped_lock lock(error_mutex_);
    if (!parked_error_) return 0;
    const size_t k = copy_message(*parked_error_, buf, cap);
    parked_error_.reset();
    return k;
}

// ---------------------------------------------------------------- image files

namespace {
pt_status image_status(ImageStatus s) {
    switch (s) {
        case ImageStatus::Ok: return PT_OK;
        case ImageStatus::InvalidArgument: return PT_ERROR_INVALID_ARGUMENT;
        case ImageStatus::Unsupported: return PT_ERROR_UNSUPPORTED;
        default: return PT_ERROR_IO;
    }
}
}  // namespace

pt_status Renderer::save_frame_png(const std::string& path) {
    if (frame_locked_here()) return fail(PT_ERROR_INVALID_STATE, "this thread holds the frame lock (pt_frame_lock)");
    std::vector<uint32_t> px;
    {
        std::scoped_lock lock(frame_mutex_);
        if (h_frame_[h_front_]) px.assign(h_frame_[h_front_], h_frame_[h_front_] + (size_t)width_ * height_);
    }
    std::string err;
    const ImageStatus s = write_png_rgba8(path, width_, height_, px.data(), true, err);
    return s == ImageStatus::Ok ? PT_OK : fail(image_status(s), err);
}

pt_status Renderer::SaveFrame(const char* path) {
    const std::string p = path ? path : "";
    ImageKind kind;
    std::string err;
    ImageStatus s = image_kind(p, width_, height_, kind, err);
    if (s != ImageStatus::Ok) return fail(image_status(s), err);
    if (kind == ImageKind::Png) return save_frame_png(p);
    std::vector<float> hdr(4 * (size_t)width_ * height_);
    const pt_status c = CopyFrameHdr(hdr.data(), (size_t)width_ * height_);
    if (c != PT_OK) return c;
    s = write_exr_rgba32f(p, width_, height_, hdr.data(), true, err);
    return s == ImageStatus::Ok ? PT_OK : fail(image_status(s), err);
}

// ---------------------------------------------------------------- hooks

pt_status Renderer::TraceBatch(const pt_ray* rays, uint32_t n, bool any, bool device, void* out, uint32_t* nodes,
                               uint32_t* tris, hipStream_t stream) {
    if (!has_scene()) return fail(PT_ERROR_INVALID_STATE, "scene buffers are not initialized");
    if (IsRendering()) return fail(PT_ERROR_INVALID_STATE, "Renderer is already rendering");
    if (n == 0) return PT_OK;
    if (!rays || !out) return fail(PT_ERROR_INVALID_ARGUMENT, "rays and output must not be NULL");
    hipStream_t s = stream ? stream : stream_;
    const size_t in_bytes = (size_t)n * sizeof(pt_ray);
    const size_t out_elem = any ? sizeof(uint32_t) : sizeof(pt_hit);
    const float4* d_in;
    void* d_out;
    if (device) {
        d_in = reinterpret_cast<const float4*>(rays);
        d_out = out;
    } else {
        PT_TRY_HIP(d_hook_in_.allocate(in_bytes), "hook input");
        PT_TRY_HIP(d_hook_out_.allocate((size_t)n * out_elem), "hook output");
        PT_TRY_HIP(hipMemcpyAsync(d_hook_in_.as<void>(), rays, in_bytes, hipMemcpyHostToDevice, s), "hook input");
        d_in = d_hook_in_.as<float4>();
        d_out = d_hook_out_.as<void>();
    }
    uint32_t* d_nodes = nullptr;
    uint32_t* d_tris = nullptr;
    if (nodes || tris) {
        PT_TRY_HIP(d_hook_aux_.allocate((size_t)n * 2 * sizeof(uint32_t)), "hook counts");
        d_nodes = d_hook_aux_.as<uint32_t>();
        d_tris = d_nodes + n;
    }
    // the overflow flag is cleared here and read by the synchronous forms; the
    // asynchronous device form leaves it set for the next synchronous call
    PT_TRY_HIP(launch_zero_u32(s, d_work_.as<unsigned int>(), 1), "zero");
    if (!device || nodes || tris) PT_TRY_HIP(launch_zero_u32(s, d_overflow_.as<unsigned int>(), 1), "zero");
    PT_TRY_HIP(launch_trace_batch(s, scene_dev(), d_in, n, any ? nullptr : static_cast<float4*>(d_out),
                                  any ? static_cast<uint32_t*>(d_out) : nullptr, d_nodes, d_tris,
                                  d_work_.as<unsigned int>(), any, hook_trace_),
               "trace launch");
    if (!device) PT_TRY_HIP(hipMemcpyAsync(out, d_out, (size_t)n * out_elem, hipMemcpyDeviceToHost, s), "hook output");
    if (nodes) PT_TRY_HIP(hipMemcpyAsync(nodes, d_nodes, (size_t)n * 4, hipMemcpyDeviceToHost, s), "hook counts");
    if (tris) PT_TRY_HIP(hipMemcpyAsync(tris, d_tris, (size_t)n * 4, hipMemcpyDeviceToHost, s), "hook counts");
    if (!device || nodes || tris) {
        uint32_t overflow = 0;
        PT_TRY_HIP(hipMemcpyAsync(&overflow, d_overflow_.as<void>(), 4, hipMemcpyDeviceToHost, s), "overflow flag");
        PT_TRY_HIP(hipStreamSynchronize(s), "hook sync");
        if (overflow) return fail(PT_ERROR_DEVICE, "traversal stack overflow");
    }
    return PT_OK;
}

// The closest-hit trace of caller device buffers with no allocation: its work
// counter, overflow flag and stack spill area live in the static scratch
// array, and the persistent grid shrinks until the spill area of the deepest
// resident BVH fits it.  The stack keeps its full depth (the spill depth is the
// tree's worst case), so it cannot overflow.  Every call on one GPU shares the
// array: calls that overlap on one GPU race.
pt_status Renderer::TraceClosestNoAlloc(const pt_ray* rays, uint32_t n, pt_hit* hits, hipStream_t stream) {
    if (!has_scene()) return fail(PT_ERROR_INVALID_STATE, "scene buffers are not initialized");
    if (IsRendering()) return fail(PT_ERROR_INVALID_STATE, "Renderer is already rendering");
    if (n == 0) return PT_OK;
    if (!rays || !hits) return fail(PT_ERROR_INVALID_ARGUMENT, "rays and hits must not be NULL");
    if (!noalloc_scratch_) return fail(PT_ERROR_DEVICE, "no scratch address on this device");
    TraceConfig cfg = hook_trace_;
    cfg.stack_cap = cfg.lds_stack + cfg.spill_depth;
    const size_t per_block = (size_t)cfg.spill_depth * cfg.block * sizeof(uint2);
    const size_t room = kNoAllocScratchBytes - kNoAllocSpillOffset;
    if (per_block > room)
        return fail(PT_ERROR_OUT_OF_RANGE, "the BVH's traversal stack needs " + std::to_string(per_block) +
                                               " bytes of scratch per block; the no-alloc trace holds " +
                                               std::to_string(room));
    if (per_block) cfg.blocks = (uint32_t)std::max<size_t>(1, std::min<size_t>(cfg.blocks, room / per_block));
    unsigned int* work = reinterpret_cast<unsigned int*>(noalloc_scratch_);
    cfg.overflow = work + 1;
    cfg.spill = reinterpret_cast<uint2*>(noalloc_scratch_ + kNoAllocSpillOffset);
    hipStream_t s = stream ? stream : stream_;
    // the counter and the flag start at zero: a fill kernel (no memset node in a captured graph)
    PT_TRY_HIP(launch_zero_u32(s, work, 2), "zero");
    PT_TRY_HIP(launch_trace_batch(s, scene_dev(), reinterpret_cast<const float4*>(rays), n,
                                  reinterpret_cast<float4*>(hits), nullptr, nullptr, nullptr, work, false, cfg),
               "trace launch");
    return PT_OK;
}

pt_status Renderer::Classify(const pt_ray* rays, uint32_t n, uint32_t* queue_of_ray, uint32_t* counts) {
    if (!has_scene()) return fail(PT_ERROR_INVALID_STATE, "scene buffers are not initialized");
    if (IsRendering()) return fail(PT_ERROR_INVALID_STATE, "Renderer is already rendering");
    if (!rays || !queue_of_ray || !counts) return fail(PT_ERROR_INVALID_ARGUMENT, "NULL argument");
    for (int q = 0; q < kQueueSlots; q++) counts[q] = 0;
    if (n == 0) return PT_OK;
    PT_TRY_HIP(d_hook_in_.allocate((size_t)n * sizeof(pt_ray)), "hook input");
    PT_TRY_HIP(d_hook_q_.allocate((size_t)n * kQueueSlots * sizeof(uint32_t) + 64), "hook queues");
    PT_TRY_HIP(d_hook_aux_.allocate((size_t)n), "hook queue ids");
    PT_TRY_HIP(hipMemcpyAsync(d_hook_in_.as<void>(), rays, (size_t)n * sizeof(pt_ray), hipMemcpyHostToDevice, stream_), "hook input");
    uint32_t* qbase = d_hook_q_.as<uint32_t>();
    uint32_t* queues[kQueueSlots];
    for (int q = 0; q < kQueueSlots; q++) queues[q] = qbase + (size_t)q * n;
    unsigned int* qcounts = reinterpret_cast<unsigned int*>(qbase + (size_t)kQueueSlots * n);
    PT_TRY_HIP(launch_zero_u32(stream_, d_work_.as<unsigned int>(), 1), "zero");
    PT_TRY_HIP(launch_zero_u32(stream_, qcounts, kQueueSlots), "zero");
    PT_TRY_HIP(launch_zero_u32(stream_, d_overflow_.as<unsigned int>(), 1), "zero");
    PT_TRY_HIP(launch_classify(stream_, scene_dev(), d_hook_in_.as<float4>(), n, d_hook_aux_.as<uint8_t>(), queues,
                               qcounts, d_work_.as<unsigned int>(), hook_trace_, bin_grid()),
               "classify launch");
    std::vector<uint32_t> host((size_t)n * kQueueSlots + 16);
    uint32_t overflow = 0;
    PT_TRY_HIP(hipMemcpyAsync(host.data(), qbase, ((size_t)n * kQueueSlots + kQueueSlots) * sizeof(uint32_t), hipMemcpyDeviceToHost, stream_), "hook queues");
    PT_TRY_HIP(hipMemcpyAsync(&overflow, d_overflow_.as<void>(), 4, hipMemcpyDeviceToHost, stream_), "overflow flag");
    PT_TRY_HIP(hipStreamSynchronize(stream_), "hook sync");
    if (overflow) return fail(PT_ERROR_DEVICE, "traversal stack overflow");
    std::vector<uint8_t> seen(n, 0);
    for (uint32_t i = 0; i < n; i++) queue_of_ray[i] = PT_INVALID_INDEX;
    uint64_t total = 0;
    for (int q = 0; q < kQueueSlots; q++) {
        const uint32_t cnt = host[(size_t)kQueueSlots * n + q];
        counts[q] = cnt;
        total += cnt;
        if (cnt > n) return fail(PT_ERROR_DEVICE, "queue count exceeds the batch");
        for (uint32_t k = 0; k < cnt; k++) {
            const uint32_t ray = host[(size_t)q * n + k];
            if (ray >= n || seen[ray]) return fail(PT_ERROR_DEVICE, "queue holds a ray twice or out of range");
            seen[ray] = 1;
            queue_of_ray[ray] = (uint32_t)q;
        }
    }
    if (total != n) return fail(PT_ERROR_DEVICE, "queues do not partition the batch");
    return PT_OK;
}

pt_status Renderer::AliasSample(const float* u, uint32_t n, uint32_t* index, float* pmf) {
    if (!has_scene()) return fail(PT_ERROR_INVALID_STATE, "scene buffers are not initialized");
    if (host0().lights.empty()) return fail(PT_ERROR_INVALID_STATE, "scene has no emissive triangles");
    if (n == 0) return PT_OK;
    if (!u || !index || !pmf) return fail(PT_ERROR_INVALID_ARGUMENT, "NULL argument");
    PT_TRY_HIP(d_hook_in_.allocate((size_t)n * 4), "hook input");
    PT_TRY_HIP(d_hook_out_.allocate((size_t)n * 8), "hook output");
    PT_TRY_HIP(hipMemcpyAsync(d_hook_in_.as<void>(), u, (size_t)n * 4, hipMemcpyHostToDevice, stream_), "hook input");
    uint32_t* d_idx = d_hook_out_.as<uint32_t>();
    float* d_pmf = reinterpret_cast<float*>(d_idx + n);
    PT_TRY_HIP(launch_alias_sample(stream_, scene_dev(), d_hook_in_.as<float>(), n, d_idx, d_pmf), "alias launch");
    PT_TRY_HIP(hipMemcpyAsync(index, d_idx, (size_t)n * 4, hipMemcpyDeviceToHost, stream_), "hook output");
    PT_TRY_HIP(hipMemcpyAsync(pmf, d_pmf, (size_t)n * 4, hipMemcpyDeviceToHost, stream_), "hook output");
    PT_TRY_HIP(hipStreamSynchronize(stream_), "hook sync");
    return PT_OK;
}

pt_status Renderer::GetAliasTable(pt_alias_bin* bins, uint32_t cap, uint32_t* count) {
    if (!has_scene()) return fail(PT_ERROR_INVALID_STATE, "scene buffers are not initialized");
    if (count) *count = (uint32_t)host0().bins.size();
    if (!bins) return PT_OK;
    if (cap < host0().bins.size()) return fail(PT_ERROR_INVALID_ARGUMENT, "capacity too small");
    for (size_t i = 0; i < host0().bins.size(); i++)
        bins[i] = pt_alias_bin{host0().bins[i].q, host0().bins[i].p, host0().bins[i].alias, host0().bins[i].value};
    return PT_OK;
}

pt_status Renderer::GetBvhInfo(pt_bvh_info& info) {
    if (!has_scene()) return fail(PT_ERROR_INVALID_STATE, "scene buffers are not initialized");
    info.node_count = host0().bvh.device_nodes ? host0().bvh.device_nodes : (uint32_t)host0().bvh.nodes.size();
    info.leaf_count = host0().bvh.leaf_count;
    info.max_depth = host0().bvh.depth;
    info.max_leaf_size = host0().bvh.max_leaf;
    info.triangle_count = host0().tri_count;
    info.light_count = (uint32_t)host0().lights.size();
    info.stack_lds = trace_.lds_stack;
    info.stack_total = trace_.lds_stack + trace_.spill_depth;
    info.node_width = host0().bvh.width;
    info.bvh8_node_count = scenes_[0]->depth8 ? (uint32_t)host0().bvh.qnodes8.size() : 0u;
    info.bvh8_depth = scenes_[0]->depth8;
    return PT_OK;
}

pt_status Renderer::GetBvh4(void* nodes, uint32_t node_cap, uint32_t* order, uint32_t order_cap) {
    if (!has_scene()) return fail(PT_ERROR_INVALID_STATE, "scene buffers are not initialized");
    if (host0().bvh.device_nodes) {
        // built on the device: read the tree back
        const ResidentScene& rs = *scenes_[0];
        const uint32_t nn = host0().bvh.device_nodes, T = host0().tri_count;
        if (nodes && node_cap < nn) return fail(PT_ERROR_INVALID_ARGUMENT, "node capacity too small");
        if (order && order_cap < T) return fail(PT_ERROR_INVALID_ARGUMENT, "order capacity too small");
        if (nodes)
            PT_TRY_HIP(hipMemcpyAsync(nodes, rs.nodes.as<void>(), (size_t)nn * sizeof(Bvh4Node), hipMemcpyDeviceToHost,
                                      stream_),
                       "BVH nodes");
        if (order)
            PT_TRY_HIP(hipMemcpyAsync(order, rs.order.as<void>(), (size_t)T * sizeof(uint32_t), hipMemcpyDeviceToHost,
                                      stream_),
                       "BVH order");
        PT_TRY_HIP(hipStreamSynchronize(stream_), "BVH read-back");
        return PT_OK;
    }
    if (nodes) {
        if (node_cap < host0().bvh.nodes.size()) return fail(PT_ERROR_INVALID_ARGUMENT, "node capacity too small");
        std::memcpy(nodes, host0().bvh.nodes.data(), host0().bvh.nodes.size() * sizeof(Bvh4Node));
    }
    if (order) {
        if (order_cap < host0().bvh.order.size()) return fail(PT_ERROR_INVALID_ARGUMENT, "order capacity too small");
        std::memcpy(order, host0().bvh.order.data(), host0().bvh.order.size() * sizeof(uint32_t));
    }
    return PT_OK;
}

pt_status Renderer::GetBvh8Quantized(void* nodes, uint32_t node_cap, uint32_t& count, uint32_t* prim_order) {
    if (!has_scene()) return fail(PT_ERROR_INVALID_STATE, "scene buffers are not initialized");
    const ResidentScene& rs = *scenes_[0];
    count = (uint32_t)rs.host.bvh.qnodes8.size();
    if (!rs.depth8) return fail(PT_ERROR_INVALID_STATE, "scene 0 was uploaded without 8-wide quantized nodes");
    if (prim_order) std::copy(rs.host.bvh.order8.begin(), rs.host.bvh.order8.end(), prim_order);
    if (!nodes) return PT_OK;
    if (node_cap < count) return fail(PT_ERROR_INVALID_ARGUMENT, "node capacity too small");
    PT_TRY_HIP(hipMemcpyAsync(nodes, rs.qnodes8.as<void>(), (size_t)count * sizeof(Bvh8QNode), hipMemcpyDeviceToHost,
                              stream_), "8-wide quantized BVH nodes");
    PT_TRY_HIP(hipStreamSynchronize(stream_), "BVH read-back");
    return PT_OK;
}

pt_status Renderer::GetBvh4Quantized(void* nodes, uint32_t node_cap) {
    if (!has_scene()) return fail(PT_ERROR_INVALID_STATE, "scene buffers are not initialized");
    const ResidentScene& rs = *scenes_[0];
    if (!rs.quantized) return fail(PT_ERROR_INVALID_STATE, "scene 0 was uploaded without quantized nodes");
    const size_t nn = rs.host.bvh.device_nodes ? rs.host.bvh.device_nodes : rs.host.bvh.nodes.size();
    if (!nodes || node_cap < nn) return fail(PT_ERROR_INVALID_ARGUMENT, "node capacity too small");
    PT_TRY_HIP(hipMemcpyAsync(nodes, rs.qnodes.as<void>(), nn * sizeof(Bvh4QNode), hipMemcpyDeviceToHost, stream_),
               "quantized BVH nodes");
    PT_TRY_HIP(hipStreamSynchronize(stream_), "BVH read-back");
    return PT_OK;
}

pt_status Renderer::GetQueueOccupancy(pt_queue_occupancy* out, uint32_t cap, uint32_t& count) const {
    count = (uint32_t)queue_history_.size();
    if (out) std::memcpy(out, queue_history_.data(), std::min<size_t>(cap, queue_history_.size()) * sizeof(*out));
    return PT_OK;
}

pt_status Renderer::GetTuning(pt_tuning& t) const {
    t = tuning_;
    return PT_OK;
}

pt_status Renderer::SetTuning(const pt_tuning& t) {
    if (IsRendering()) return fail(PT_ERROR_INVALID_STATE, "Renderer is already rendering");
    if (!trace_config_supported(t.trace_block, t.trace_block >= 256 ? 8 : 16))
        return fail(PT_ERROR_INVALID_ARGUMENT, "trace_block must be 64, 128, 256 or 512");
    if (t.bvh_width != 2 && t.bvh_width != 4) return fail(PT_ERROR_INVALID_ARGUMENT, "bvh_width must be 2 or 4");
    if (t.bvh_builder != PT_BVH_BUILDER_HOST && t.bvh_builder != PT_BVH_BUILDER_GPU)
        return fail(PT_ERROR_INVALID_ARGUMENT, "bvh_builder must be PT_BVH_BUILDER_HOST or PT_BVH_BUILDER_GPU");
    if (t.bvh_builder == PT_BVH_BUILDER_GPU && t.bvh_width != 4)
        return fail(PT_ERROR_INVALID_ARGUMENT, "the GPU BVH builder emits 4-wide nodes");
    if (t.bvh_quantized > 15) return fail(PT_ERROR_INVALID_ARGUMENT, "bvh_quantized must be in [0, 15]");
    if ((t.bvh_quantized & 8u) && !(t.bvh_quantized & 4u))
        return fail(PT_ERROR_INVALID_ARGUMENT, "bvh_quantized bit 3 (closest hit on the 8-wide tree) needs bit 2");
    if (t.shade_block != 256) return fail(PT_ERROR_INVALID_ARGUMENT, "shade_block must be 256");
    if (t.trace_blocks_per_cu > 64 || t.shadow_blocks_per_cu > 64)
        return fail(PT_ERROR_INVALID_ARGUMENT, "traversal blocks per CU must be at most 64");
    if (t.pool_size != 0 && (t.pool_size < 1024 || t.pool_size > kAutoPoolMax))
        return fail(PT_ERROR_INVALID_ARGUMENT, "pool_size must be 0 (automatic) or in [1024, 2^27]");
    if (t.trace_refill < 1 || t.trace_refill > 64) return fail(PT_ERROR_INVALID_ARGUMENT, "trace_refill must be in [1, 64]");
    if (t.bvh_max_leaf < 1 || t.bvh_max_leaf > kMaxLeafSize) return fail(PT_ERROR_INVALID_ARGUMENT, "bvh_max_leaf must be in [1, 16]");
    if (!(t.bvh_traversal_cost > 0.0f) || t.bvh_traversal_cost > 100.0f)
        return fail(PT_ERROR_INVALID_ARGUMENT, "bvh_traversal_cost must be in (0, 100]");
    tuning_ = t;
    invalidate_graph();
    if (d_hit_.bytes()) PT_TRY(ensure_pool());
    if (has_scene()) PT_TRY(configure_trace());
    return PT_OK;
}

}  // namespace pt
