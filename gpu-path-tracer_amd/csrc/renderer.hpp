// Renderer of the core: owns the device buffers and runs the wavefront loop.
// Mirrors Core::Graphics::Cuda::Renderer
// (Core/Include/Core/Graphics/Cuda/PathTracing/Renderer.hpp:42-189).
#pragma once

#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <cstdint>
#include <memory>
#include <mutex>
#include <optional>
#include <string>
#include <thread>
#include <vector>

#include "../../include/pt_core.h"
#include "launch.hpp"
#include "scene.hpp"

namespace pt {

class DeviceBuffer {
public:
    DeviceBuffer() = default;
    ~DeviceBuffer() { release(); }
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
    hipError_t allocate(size_t bytes);     // keeps the allocation if it is large enough
    void release();
    template <class T> T* as() const { return static_cast<T*>(ptr_); }
    size_t bytes() const { return bytes_; }
private:
    void* ptr_ = nullptr;
    size_t bytes_ = 0;
};

struct ServerState;   // server.hpp

// Width and height of a frame rendered with `s` (clamped as RenderFrame does).
void frame_size(const pt_render_settings& s, uint32_t& width, uint32_t& height);

// Options of one wavefront run over a set of views.
struct RunOptions {
    bool clear = true;           // zero the accumulation pixels first
    bool guided = false;         // guide MIS and training records (one view)
    float guide_fraction = 0.0f;
    uint32_t rec_shift = 0;      // training records from one sample in 2^rec_shift
};

// A scene resident on the device: host copy (for the hooks and the BVH
// queries) and its device records.
struct ResidentScene {
    HostScene host;
    DeviceBuffer nodes, tris, shade, materials, lights, bins;
    DeviceBuffer order;        // GPU-built BVH: original triangle of each BVH slot (the host copy is empty)
    DeviceBuffer qnodes;       // quantized copy of the 4-wide nodes
    DeviceBuffer qnodes8;      // the same binary tree 8-wide and quantized (any hit), or empty
    DeviceBuffer tris8;        // triangle records in the 8-wide tree's leaf order, or empty
    DeviceBuffer top[2];       // top-of-tree tables of nodes and qnodes (the 8-wide tree needs none)
    uint32_t quantized = 0;    // traversals that use qnodes: bit 0 closest hit, bit 1 any hit
    uint32_t depth8 = 0;       // levels of the 8-wide tree (0: none)
    SceneDev dev() const;
};

class Renderer {
public:
    Renderer();
    ~Renderer();
    Renderer(const Renderer&) = delete;
    Renderer& operator=(const Renderer&) = delete;

    pt_status init();   // device, stream, small buffers
    pt_status InitializeRenderingBuffers(uint32_t width, uint32_t height);
    pt_status InitializeSceneBuffers(const pt_scene_desc& scene);
    pt_status Clear(uint8_t r, uint8_t g, uint8_t b, uint8_t a);
    pt_status StartSimulation(const pt_render_settings& s, const pt_camera& camera,
                              const pt_motion_state* states, uint32_t n, const char* output_folder);
    pt_status ResumeSimulation(uint32_t start_frame);
    pt_status StopRendering();
    bool IsRendering() const { return is_rendering_.load(); }
    void GetProgress(pt_progress& p) const;
    pt_status CopyFrame(uint32_t* rgba, size_t n);
    pt_status LockFrame(const uint32_t** rgba, size_t* n);
    pt_status UnlockFrame();
    // the frame as ".png" (RGBA8) or ".exr" (float radiance of the last single frame)
    pt_status SaveFrame(const char* path);
    pt_status CopyFrameHdr(float* rgba, size_t n);
    size_t PeekLastError(char* buf, size_t cap) const;
    size_t ConsumeLastError(char* buf, size_t cap);

    pt_status RenderFrameSync(const pt_render_settings& s, const pt_camera& camera, const pt_motion_state& state,
                              uint32_t flags, pt_frame_stats* stats);

    // resident scenes and batches of views
    pt_status AddScene(const pt_scene_desc& scene, uint32_t& id);
    pt_status RemoveScene(uint32_t id);
    pt_status RenderBatch(const pt_view* views, uint32_t n, float* const* hdr, uint32_t* const* rgba, uint32_t flags,
                          pt_frame_stats* stats);

    // learned path guiding (guide_host.cpp)
    pt_status GuideReset(uint64_t seed);
    pt_status GuideGetWeights(float* w, uint32_t n);
    pt_status GuideSetWeights(const float* w, uint32_t n);
    pt_status GuideSetBox(const float lo[3], const float hi[3]);
    pt_status GuideForward(const float* features, uint32_t n, float* out);
    pt_status GuideMixture(const float* features, uint32_t n, float* lobes);
    pt_status GuideGradient(const pt_guide_record* recs, uint32_t n, float w_max, float* grad, float* loss);
    pt_status GuideTrain(const pt_guide_record* recs, uint32_t n, uint32_t steps, uint32_t batch, float lr, float w_max,
                         float* losses);
    pt_status RenderGuided(const pt_render_settings& s, const pt_camera& camera, const pt_motion_state& state,
                           const pt_guide_settings& g, pt_guide_stats* gstats, pt_frame_stats* stats);

    // serving (server.cpp)
    pt_status ServerStart(const pt_server_config& cfg);
    pt_status ServerSubmit(const pt_view& view, float* hdr, uint32_t* rgba, uint64_t& id);
    pt_status ServerWait(uint64_t id, uint32_t timeout_ms, pt_request_info& info);
    pt_status ServerStats(pt_server_stats& stats);
    pt_status ServerStop(pt_server_stats* stats);

    // hooks
    pt_status TraceBatch(const pt_ray* rays, uint32_t n, bool any, bool device, void* out, uint32_t* nodes,
                         uint32_t* tris, hipStream_t stream);
    pt_status Classify(const pt_ray* rays, uint32_t n, uint32_t* queue_of_ray, uint32_t* counts);
    // TraceBatch(closest, device) with its scratch in the static device array
    // (launch.hpp): allocates nothing
    pt_status TraceClosestNoAlloc(const pt_ray* rays_device, uint32_t n, pt_hit* hits_device, hipStream_t stream);
    pt_status AliasSample(const float* u, uint32_t n, uint32_t* index, float* pmf);
    pt_status GetAliasTable(pt_alias_bin* bins, uint32_t cap, uint32_t* count);
    pt_status GetBvhInfo(pt_bvh_info& info);
    pt_status GetQueueOccupancy(pt_queue_occupancy* out, uint32_t cap, uint32_t& count) const;
    pt_status GetBvh4(void* nodes, uint32_t node_cap, uint32_t* order, uint32_t order_cap);
    pt_status GetBvh4Quantized(void* nodes, uint32_t node_cap);
    pt_status GetBvh8Quantized(void* nodes, uint32_t node_cap, uint32_t& count, uint32_t* prim_order);
    pt_status GetTuning(pt_tuning& t) const;
    pt_status SetTuning(const pt_tuning& t);

    const char* error_message() const { return last_error_.c_str(); }
    // Records the message of a failing call (the server threads fail too).
    pt_status fail(pt_status s, const std::string& msg) {
        std::scoped_lock lock(message_mutex_);
        last_error_ = msg;
        return s;
    }

private:
    // Which kernel instances a set of views needs: per-ray scene lookups when
    // the views span several scenes, per-path view lookups when there are
    // several views.  One captured graph per mode.
    struct LoopMode {
        bool trace_multi = false;
        bool shade_multi = false;
        uint32_t scene = 0;          // the one scene when !trace_multi
        bool guided = false;         // guide MIS and training records (one view)
        float guide_fraction = 0.0f; // baked into the captured kernels' arguments
        uint32_t rec_shift = 0;      // likewise
        int key() const {
            if (guided)
                return (1 << 30) | (int)(rec_shift << 20) | (int)(scene << 12) | (int)(guide_fraction * 1000.0f + 0.5f);
            return trace_multi ? -1 - (int)shade_multi : (int)(scene * 2u + (shade_multi ? 1u : 0u));
        }
    };

    struct GraphEntry {
        int key;
        hipGraph_t graph;
        hipGraphExec_t exec;
    };
    pt_status check_hip(hipError_t e, const char* what);
    pt_status ensure_pool(uint64_t samples = 0, bool exact = false);
    pt_status ensure_accum(uint32_t pixels);
    pt_status set_frame_size(uint32_t width, uint32_t height);
    pt_status publish_frame();
    pt_status configure_trace();
    pt_status upload_scene(uint32_t id, const pt_scene_desc& desc);
    pt_status make_view(const pt_view& v, ViewRecord& rec, uint32_t& depth);
    pt_status ensure_graph(const LoopMode& mode, hipGraphExec_t& exec);
    pt_status run_views(const std::vector<ViewRecord>& views, uint32_t max_depth, uint32_t flags, pt_frame_stats* stats,
                        const std::atomic<bool>* stop, const RunOptions& opt = RunOptions());
    pt_status render_frame(const pt_render_settings& s, const pt_camera& camera, const pt_motion_state& state,
                           uint32_t flags, pt_frame_stats* stats, const std::atomic<bool>* stop);
    // One bounce.  With `fork` the any-hit traversal runs on side_stream_ and
    // overlaps the regeneration, the bookkeeping and the next bounce's
    // closest-hit traversal and binning; the next shade waits for it.
    pt_status launch_iteration(int cur, bool count, const LoopMode& mode, bool fork = false);
    pt_status join_shadow();
    void server_loop();
    void readback_loop();
    pt_status server_step();
    pt_status save_frame_png(const std::string& path);
    void simulation_loop(uint32_t start_frame);
    void invalidate_graph();
    PathState path_state(const LoopMode* mode = nullptr) const;
    pt_status ensure_guide();
    pt_status ensure_guide_pool(uint32_t record_capacity);
    GuideDev guide_dev() const;
    void guide_box_from_scene();
    uint32_t bin_grid() const { return (uint32_t)std::max(1, cu_count_) * 4u; }
    bool has_scene() const { return scenes_[0] != nullptr; }
    const HostScene& host0() const { return scenes_[0]->host; }
    SceneDev scene_dev(uint32_t id = 0) const { return scenes_[id]->dev(); }

    int device_ = 0;
    int cu_count_ = 0;
    hipStream_t stream_ = nullptr;
    hipStream_t side_stream_ = nullptr;     // any-hit traversal overlapped with the next bounce (graphs)
    hipEvent_t fork_events_[2] = {}, join_events_[2] = {};
    int shadow_pending_ = -1;               // parity of a forked any-hit launch not yet joined
    std::string last_error_;
    std::mutex message_mutex_;

    // resident scenes; id 0 is the scene of InitializeSceneBuffers
    std::unique_ptr<ResidentScene> scenes_[kMaxScenes];
    DeviceBuffer d_scene_table_;            // kMaxScenes SceneDev records
    DeviceBuffer d_views_;                  // kViewRing ViewRecords

    // path pool and queues
    pt_tuning tuning_{};
    uint32_t pool_capacity_ = 0;
    DeviceBuffer d_ray_o_[2], d_ray_d_[2], d_thr_[2], d_meta_[2];
    DeviceBuffer d_hit_, d_sh_o_, d_sh_d_, d_sh_L_, d_q_[kQueueSlots], d_qid_;
    DeviceBuffer d_ctr_, d_fp_, d_work_, d_overflow_;
    FrameParams* h_fp_ = nullptr;           // pinned staging
    Counters* h_ctr_ring_ = nullptr;        // pinned, kRing entries
    std::vector<pt_queue_occupancy> queue_history_;   // iterations of the last profiled frame
    static constexpr int kRing = 4;
    static constexpr int kGraphIterations = 4;   // bounces per graph launch (even: parity returns to 0)
    hipEvent_t ring_events_[kRing] = {};

    // pixels: accumulation capacity for all views of a run; width_ x height_
    // is the frame of the single-frame API
    uint32_t width_ = 0, height_ = 0;
    uint32_t accum_capacity_ = 0;
    bool hdr_valid_ = false;                // accum[0, w*h) holds the last single frame
    DeviceBuffer d_accum_, d_frame_;
    // The last finished frame (RGBA8, width_ x height_): two pinned buffers.
    // Readers copy the front one under frame_mutex_; a finished frame is copied
    // from the device into the back one and made the front by a swap.
    uint32_t* h_frame_[2] = {nullptr, nullptr};
    size_t h_frame_capacity_ = 0;           // pixels per buffer
    int h_front_ = 0;
    mutable std::mutex frame_mutex_;
    // the thread holding frame_mutex_ through LockFrame (default id: none)
    std::atomic<std::thread::id> frame_lock_owner_{};
    bool frame_locked_here() const { return frame_lock_owner_.load() == std::this_thread::get_id(); }
    uint32_t last_spp_ = 1;

    // traversal
    TraceConfig trace_{};                   // closest hit, one scene per launch
    TraceConfig shadow_trace_{};            // any hit, with its own stack spill area
    TraceConfig trace_multi_{};             // the same two for launches over several scenes, whose
    TraceConfig shadow_trace_multi_{};      //   kernels hold more registers (grids sized to their occupancy)
    TraceConfig shadow_alone_{};            // any hit where no closest-hit launch runs beside it
    TraceConfig shadow_alone_multi_{};      //   (stage-by-stage loops): the whole occupancy
    TraceConfig hook_trace_{};              // trace_ for the caller-batch hooks, with their own overflow flag
    DeviceBuffer d_req_invalid_;            // server: per-request invalid-pixel counters
    unsigned char* noalloc_scratch_ = nullptr;  // this device's address of the no-alloc scratch array
    uint32_t* h_req_invalid_ = nullptr;     //   and their pinned host copies
    DeviceBuffer d_spill_;
    uint32_t stack_total_ = 0;

    // graphs of two bounce iterations (ping and pong), one per loop mode
    std::vector<GraphEntry> graphs_;

    // hook scratch
    DeviceBuffer d_hook_in_, d_hook_out_, d_hook_aux_, d_hook_q_;

    // async rendering (Renderer.cpp:330-466)
    std::thread render_thread_;
    std::atomic<bool> stop_requested_{false};
    std::atomic<bool> is_rendering_{false};
    std::atomic<uint32_t> done_frames_{0};
    std::atomic<uint64_t> done_samples_{0};
    uint32_t total_frames_ = 0;
    uint64_t total_samples_ = 0;
    pt_render_settings settings_{};
    pt_camera camera_{};
    std::vector<pt_motion_state> motion_;
    std::string output_folder_;
    mutable std::mutex error_mutex_;
    std::optional<std::string> parked_error_;

    // guide network and its training records
    DeviceBuffer g_w32_, g_w_, g_wt_, g_wf_, g_bf_, g_grad_, g_m_, g_v_, g_loss_, g_part_;
    DeviceBuffer g_recs_, g_head_[2], g_sh_head_, g_rr_scale_[2];
    uint32_t g_rec_capacity_ = 0;
    uint32_t g_pool_ = 0;            // path slots the per-path guide buffers cover
    uint32_t g_step_ = 0;            // Adam steps taken
    bool g_init_ = false;
    bool g_box_set_ = false;         // box given by GuideSetBox (else scene 0's bounds)
    GuideConfig g_cfg_{};

    std::unique_ptr<ServerState> server_;
    LoopMode server_mode_;
    hipGraphExec_t server_graph_ = nullptr;
    std::vector<hipGraphExec_t> server_scene_graphs_;   // by scene id: the one-scene traversal loops
};

#define PT_TRY_HIP(expr, what)                              \
    do {                                                    \
        pt_status _s = check_hip((expr), (what));           \
        if (_s != PT_OK) return _s;                         \
    } while (0)

#define PT_TRY(expr)                                        \
    do {                                                    \
        pt_status _s = (expr);                              \
        if (_s != PT_OK) return _s;                         \
    } while (0)

}  // namespace pt
