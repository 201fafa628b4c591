// Renderer of the core: owns the device buffers and runs the wavefront loop.
// Mirrors Core::Graphics::Cuda::Renderer
// (Core/Include/Core/Graphics/Cuda/PathTracing/Renderer.hpp:42-189).
#pragma once

#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <cstdint>
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
    pt_status CopyFrameHdr(float* rgba, size_t n);
    size_t PeekLastError(char* buf, size_t cap) const;
    size_t ConsumeLastError(char* buf, size_t cap);

    pt_status RenderFrameSync(const pt_render_settings& s, const pt_camera& camera, const pt_motion_state& state,
                              uint32_t flags, pt_frame_stats* stats);

    // hooks
    pt_status TraceBatch(const pt_ray* rays, uint32_t n, bool any, bool device, void* out, uint32_t* nodes,
                         uint32_t* tris, hipStream_t stream);
    pt_status Classify(const pt_ray* rays, uint32_t n, uint32_t* queue_of_ray, uint32_t* counts);
    pt_status AliasSample(const float* u, uint32_t n, uint32_t* index, float* pmf);
    pt_status GetAliasTable(pt_alias_bin* bins, uint32_t cap, uint32_t* count);
    pt_status GetBvhInfo(pt_bvh_info& info);
    pt_status GetBvh4(void* nodes, uint32_t node_cap, uint32_t* order, uint32_t order_cap);
    pt_status GetTuning(pt_tuning& t) const;
    pt_status SetTuning(const pt_tuning& t);

    const char* error_message() const { return last_error_.c_str(); }
    pt_status fail(pt_status s, const std::string& msg) {
        last_error_ = msg;
        return s;
    }

private:
    struct FrameSetup {
        uint32_t width, height, spp, axis, depth;
        bool nee;
        uint64_t total;
    };
    pt_status check_hip(hipError_t e, const char* what);
    pt_status ensure_pool();
    pt_status ensure_pixels(uint32_t width, uint32_t height);
    pt_status configure_trace();
    pt_status ensure_graph();
    pt_status render_frame(const pt_render_settings& s, const pt_camera& camera, const pt_motion_state& state,
                           uint32_t flags, pt_frame_stats* stats, const std::atomic<bool>* stop);
    pt_status launch_iteration(int cur, bool count);
    pt_status save_frame_png(const std::string& path);
    void simulation_loop(uint32_t start_frame);
    void invalidate_graph();
    PathState path_state() const;
    uint32_t bin_grid() const { return (uint32_t)std::max(1, cu_count_) * 4u; }
    SceneDev scene_dev() const;

    int device_ = 0;
    int cu_count_ = 0;
    hipStream_t stream_ = nullptr;
    std::string last_error_;

    // scene
    bool has_scene_ = false;
    HostScene host_;
    DeviceBuffer d_nodes_, d_tris_, d_shade_, d_materials_, d_lights_, d_bins_;

    // path pool and queues
    pt_tuning tuning_{};
    uint32_t pool_capacity_ = 0;
    DeviceBuffer d_ray_o_[2], d_ray_d_[2], d_thr_[2], d_meta_[2];
    DeviceBuffer d_hit_, d_sh_o_, d_sh_d_, d_sh_L_, d_q_[kQueueSlots], d_qid_;
    DeviceBuffer d_ctr_, d_fp_, d_work_, d_overflow_;
    FrameParams* h_fp_ = nullptr;           // pinned staging
    Counters* h_ctr_ring_ = nullptr;        // pinned, kRing entries
    static constexpr int kRing = 4;
    hipEvent_t ring_events_[kRing] = {};

    // pixels
    uint32_t width_ = 0, height_ = 0;
    DeviceBuffer d_accum_, d_frame_;
    uint32_t* h_frame_staging_ = nullptr;   // pinned
    size_t h_frame_staging_pixels_ = 0;
    std::vector<uint32_t> frame_;           // last finished frame, guarded by frame_mutex_
    mutable std::mutex frame_mutex_;
    uint32_t last_spp_ = 1;

    // traversal
    TraceConfig trace_{};
    DeviceBuffer d_spill_;
    uint32_t stack_total_ = 0;

    // graph of two bounce iterations (ping and pong)
    hipGraphExec_t graph_exec_ = nullptr;
    hipGraph_t graph_ = nullptr;

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
};

}  // namespace pt
