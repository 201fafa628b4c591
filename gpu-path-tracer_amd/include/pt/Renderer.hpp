// C++ facade over include/pt_core.h with the method names of
// Core::Graphics::Cuda::Renderer (Core/Include/Core/Graphics/Cuda/PathTracing/Renderer.hpp:42-189).
//
// Header-only; links against libptcore.so.  Where the reference returns
// std::expected<void, Core::Utils::Error>, these methods return pt::Result,
// which is false on failure and carries the status and the library's message.
#pragma once

#include <cstdint>
#include <optional>
#include <string>
#include <utility>
#include <vector>

#include "../../../include/pt_core.h"

namespace pt {

struct Result {
    pt_status status = PT_OK;
    std::string message;
    explicit operator bool() const { return status == PT_OK; }
};

class Renderer {
public:
    // FrameView::Lock() -> LockedFrameView (Renderer.hpp:45-104): the published
    // frame itself, held under the renderer's frame lock until the view is
    // destroyed (pt_frame_lock / pt_frame_unlock). Destroy it on the thread that
    // made it; while it lives, that thread's frame-taking calls fail.
    class LockedFrameView {
    public:
        LockedFrameView(LockedFrameView&& o) noexcept
            : h_(std::exchange(o.h_, nullptr)), data_(o.data_), size_(o.size_) {}
        LockedFrameView& operator=(LockedFrameView&&) = delete;
        LockedFrameView(const LockedFrameView&) = delete;
        LockedFrameView& operator=(const LockedFrameView&) = delete;
        ~LockedFrameView() {
            if (h_) pt_frame_unlock(h_);
        }
        const uint32_t* GetData() const { return data_; }
        size_t GetSize() const { return size_; }
        size_t GetElementSize() const { return sizeof(uint32_t); }
    private:
        friend class Renderer;
        explicit LockedFrameView(pt_renderer* h) {
            if (pt_frame_lock(h, &data_, &size_) == PT_OK) h_ = h;
            else { data_ = nullptr; size_ = 0; }
        }
        pt_renderer* h_ = nullptr;
        const uint32_t* data_ = nullptr;
        size_t size_ = 0;
    };

    class FrameView {
    public:
        explicit FrameView(const Renderer& r) : r_(&r) {}
        size_t GetSize() const { return (size_t)r_->GetPixelGridWidth() * r_->GetPixelGridHeight(); }
        size_t GetElementSize() const { return sizeof(uint32_t); }
        LockedFrameView Lock() const { return LockedFrameView(r_->h_); }
    private:
        const Renderer* r_;
    };

    Renderer() { created_ = make_result(pt_renderer_create(&h_)); }
    ~Renderer() {
        if (h_) pt_renderer_destroy(h_);
    }
    Renderer(const Renderer&) = delete;
    Renderer& operator=(const Renderer&) = delete;
    Renderer(Renderer&&) = delete;
    Renderer& operator=(Renderer&&) = delete;

    // Status of construction: a renderer needs a visible GPU.
    const Result& Created() const { return created_; }

    Result InitializeRenderingBuffers(uint32_t width, uint32_t height) {
        return make_result(pt_initialize_rendering_buffers(h_, width, height));
    }
    // Replaces the (entt::registry, Assets::Storage) pair with flattened world-space arrays.
    Result InitializeSceneBuffers(const pt_scene_desc& scene) {
        return make_result(pt_initialize_scene_buffers(h_, &scene));
    }
    Result Clear(uint8_t r, uint8_t g, uint8_t b, uint8_t a) { return make_result(pt_clear(h_, r, g, b, a)); }
    Result StartSimulation(uint32_t width, uint32_t height, const pt_camera& camera,
                           std::vector<pt_motion_state> cameraMotionStates, uint32_t samplesPerPixel,
                           uint32_t pathDepth, const std::string& outputFolder, bool useNextEventEstimation = true,
                           uint64_t seed = 1) {
        motion_ = std::move(cameraMotionStates);
        pt_render_settings s{width, height, samplesPerPixel, pathDepth, useNextEventEstimation ? 1u : 0u, 0u, seed};
        return make_result(pt_start_simulation(h_, &s, &camera, motion_.data(), (uint32_t)motion_.size(),
                                               outputFolder.empty() ? nullptr : outputFolder.c_str()));
    }
    Result ResumeSimulation(uint32_t startFrame) { return make_result(pt_resume_simulation(h_, startFrame)); }
    Result StopRendering() { return make_result(pt_stop_rendering(h_)); }

    FrameView GetFrameView() const { return FrameView(*this); }
    bool IsRendering() const { return pt_is_rendering(h_) != 0; }
    uint32_t GetPixelGridWidth() const { return progress().width; }
    uint32_t GetPixelGridHeight() const { return progress().height; }
    uint32_t GetDoneFrames() const { return progress().done_frames; }
    uint32_t GetTotalFrames() const { return progress().total_frames; }
    uint64_t GetDoneSamples() const { return progress().done_samples; }
    uint64_t GetTotalSamples() const { return progress().total_samples; }
    uint32_t GetSamplesPerPixel() const { return progress().samples_per_pixel; }
    uint32_t GetSamplesPerPixelAxis() const { return progress().samples_per_pixel_axis; }
    uint32_t GetPathDepth() const { return progress().path_depth; }

    std::optional<std::string> PeekLastError() const {
        char buf[4096];
        return pt_peek_last_error(h_, buf, sizeof(buf)) ? std::optional<std::string>(buf) : std::nullopt;
    }
    std::optional<std::string> ConsumeLastError() {
        char buf[4096];
        return pt_consume_last_error(h_, buf, sizeof(buf)) ? std::optional<std::string>(buf) : std::nullopt;
    }

    // Synchronous frame (no reference counterpart: RenderFrame is private there).
    Result RenderFrame(const pt_render_settings& s, const pt_camera& camera, const pt_motion_state& state,
                       pt_frame_stats* stats = nullptr, uint32_t flags = PT_FRAME_DEFAULT) {
        return make_result(pt_render_frame(h_, &s, &camera, &state, flags, stats));
    }

    pt_renderer* handle() const { return h_; }

private:
    Result make_result(pt_status s) const {
        Result r;
        r.status = s;
        if (s != PT_OK) r.message = h_ ? pt_get_error_message(h_) : "renderer could not be created";
        return r;
    }
    pt_progress progress() const {
        pt_progress p{};
        pt_get_progress(h_, &p);
        return p;
    }

    pt_renderer* h_ = nullptr;
    Result created_;
    std::vector<pt_motion_state> motion_;
};

}  // namespace pt
