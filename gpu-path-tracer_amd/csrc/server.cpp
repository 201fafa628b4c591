// Serving loop: one continuous wavefront over every admitted request.
//
// The reference serves one request at a time: StartSimulation hands a render
// thread one camera path and RenderFrame runs one frame's bounce loop to the
// end (Renderer.cpp:330-399,468-647).  Here a server thread keeps the bounce
// loop running while requests come and go:
//
//  * admission: a request is admitted when its pixels fit the accumulation
//    arena (the memory budget) and the admitted count is below the cap;
//  * scheduling: admitted requests are cut into chunks of samples, each a
//    record of the device view ring; the thread keeps at most `window`
//    unlaunched samples on the device and fills it with the request that has
//    the fewest samples left, so small requests do not wait behind large ones;
//  * completion: every iteration moves every live path one bounce, so a path
//    launched in iteration i has ended, shadow ray included, after iteration
//    i + depth_limit.  A request whose last sample was launched by iteration
//    count L is finished once L + depth_limit iterations have run: no per-path
//    bookkeeping on the device;
//  * readback: a second thread copies finished pixels to the caller and frees
//    the arena range, so the launch pipeline never waits on a copy.
#include <algorithm>
#include <cstddef>
#include <cstring>

#include "renderer.hpp"
#include "server.hpp"

namespace pt {

namespace {
constexpr uint64_t kDefaultBudget = 2ull << 30;
constexpr uint32_t kDefaultMaxRequests = 1024;
constexpr uint32_t kMaxRequests = 2048;
// invalid-pixel counters, each held by one request from its postprocess to
// its readback (the server thread waits for one if all are held)
constexpr uint32_t kInvalidSlots = 4096;
constexpr uint64_t kPixelBytes = sizeof(float4) + sizeof(uint32_t);
constexpr uint64_t kServerAutoPool = 1ull << 25;   // path slots of a server started with an automatic pool
}  // namespace

pt_status Renderer::ServerStart(const pt_server_config& cfg_in) {
    if (server_) return fail(PT_ERROR_INVALID_STATE, "the server is already running");
    if (IsRendering()) return fail(PT_ERROR_INVALID_STATE, "Renderer is already rendering");
    uint32_t first_scene = kMaxScenes;
    for (uint32_t k = 0; k < kMaxScenes && first_scene == kMaxScenes; k++)
        if (scenes_[k]) first_scene = k;
    if (first_scene == kMaxScenes) return fail(PT_ERROR_INVALID_STATE, "no resident scene");
    for (uint32_t k = 0; k < kMaxScenes; k++)
        if (scenes_[k] && (scenes_[k]->host.bvh.width != scenes_[first_scene]->host.bvh.width ||
                           scenes_[k]->quantized != scenes_[first_scene]->quantized ||
                           (scenes_[k]->depth8 != 0) != (scenes_[first_scene]->depth8 != 0)))
            return fail(PT_ERROR_INVALID_STATE, "the resident scenes must share one BVH node format");
    pt_server_config cfg = cfg_in;
    if (cfg.max_requests == 0) cfg.max_requests = kDefaultMaxRequests;
    if (cfg.max_requests > kMaxRequests)
        return fail(PT_ERROR_INVALID_ARGUMENT, "max_requests must be at most " + std::to_string(kMaxRequests));
    if (cfg.accum_budget_bytes == 0) cfg.accum_budget_bytes = kDefaultBudget;
    const uint64_t budget_pixels = std::min<uint64_t>(cfg.accum_budget_bytes / kPixelBytes, 0x7FFFFFFFull);
    if (budget_pixels < 1) return fail(PT_ERROR_INVALID_ARGUMENT, "accum_budget_bytes holds no pixel");
    // requests are not known yet: an automatic pool is set to 2^25 slots with a
    // window of four pools (eight rooms of 1-20M triangles: 2^27 slots and
    // their four-pool window gave 0.96x of one request at a time against
    // 0.97x, and p50 474 / 972 ms against 360 / 414 ms at 50 / 80 % load)
    PT_TRY(ensure_pool(kServerAutoPool, true));
    // the window covers what the wavefront regenerates while the host reads its
    // counters a graph launch late (about a pool per three bounces): with two
    // pools the path pool ran part-empty (eight rooms of 1-20M triangles: 0.97x
    // the rays/s of one request at a time; four pools 1.00x).  A deeper window
    // commits samples further ahead of newly arriving requests: eight pools
    // gave the same rays/s at a higher median latency
    if (cfg.window_samples == 0) cfg.window_samples = 4ull * pool_capacity_;
    if (cfg.chunk_samples == 0) cfg.chunk_samples = std::max<uint64_t>(pool_capacity_ / 4, 4096);
    PT_TRY(ensure_accum((uint32_t)budget_pixels));
    hdr_valid_ = false;

    if (!h_req_invalid_) {
        PT_TRY_HIP(d_req_invalid_.allocate(kInvalidSlots * sizeof(uint32_t)), "request counters");
        PT_TRY_HIP(hipHostMalloc((void**)&h_req_invalid_, kInvalidSlots * sizeof(uint32_t), hipHostMallocDefault),
                   "hipHostMalloc");
    }
    auto S = std::make_unique<ServerState>();
    S->cfg = cfg;
    S->arena.reset((uint32_t)budget_pixels);
    for (uint32_t k = kInvalidSlots; k-- > 0;) S->free_slots.push_back(k);
    PT_TRY_HIP(hipStreamCreateWithFlags(&S->rb_stream, hipStreamNonBlocking), "hipStreamCreate");

    FrameParams& fp = *h_fp_;
    std::memset(&fp, 0, sizeof(fp));
    for (uint32_t b = 0; b <= kMaxDepthLimit; b++) fp.bounce_jump[b] = pcg_jump(bounce_draw_base(b));
    const PathState p = path_state();
    hipError_t e = launch_write_words(stream_, d_fp_.as<void>(), h_fp_, sizeof(FrameParams));
    if (e == hipSuccess) e = launch_frame_begin(stream_, p);
    if (e == hipSuccess) e = hipStreamSynchronize(stream_);
    if (e != hipSuccess) {
        (void)hipStreamDestroy(S->rb_stream);
        return check_hip(e, "server start");
    }
    server_mode_ = LoopMode{};
    server_mode_.trace_multi = true;
    server_mode_.shade_multi = true;
    server_mode_.scene = first_scene;
    server_graph_ = nullptr;
    server_scene_graphs_.assign(kMaxScenes, nullptr);
    if (tuning_.use_graph) {
        pt_status s = ensure_graph(server_mode_, server_graph_);
        // the one-scene loops of every resident scene, captured now: a capture
        // while the server threads run can be invalidated by the read-back
        // thread's copies
        for (uint32_t id = 0; id < kMaxScenes && s == PT_OK; id++) {
            if (!scenes_[id]) continue;
            LoopMode m = server_mode_;
            m.trace_multi = false;
            m.scene = id;
            s = ensure_graph(m, server_scene_graphs_[id]);
        }
        if (s != PT_OK) {
            (void)hipStreamDestroy(S->rb_stream);
            return s;
        }
    }
    S->t0 = std::chrono::steady_clock::now();
    server_ = std::move(S);
    is_rendering_.store(true);
    try {
        server_->readback = std::thread([this]() { readback_loop(); });
        server_->thread = std::thread([this]() { server_loop(); });
    } catch (...) {
        {
            std::scoped_lock lock(server_->mu);
            server_->stop = true;
            server_->loop_exited = true;
        }
        server_->cv_rb.notify_all();
        if (server_->readback.joinable()) server_->readback.join();
        (void)hipStreamDestroy(server_->rb_stream);
        server_.reset();
        is_rendering_.store(false);
        return fail(PT_ERROR_INVALID_STATE, "failed to start the server threads");
    }
    return PT_OK;
}

pt_status Renderer::ServerSubmit(const pt_view& view, float* hdr, uint32_t* rgba, uint64_t& id) {
    if (!server_) return fail(PT_ERROR_INVALID_STATE, "the server is not running");
    if (view.scene >= kMaxScenes || !scenes_[view.scene])
        return fail(PT_ERROR_INVALID_ARGUMENT, "view names scene " + std::to_string(view.scene) + ", which is not resident");
    auto r = std::make_shared<ServerRequest>();
    PT_TRY(make_view(view, r->rec, r->depth));
    r->hdr = hdr;
    r->rgba = rgba;
    r->total = r->rec.sample_count;
    r->pixels = r->rec.width * r->rec.height;
    ServerState& S = *server_;
    if (r->pixels > S.arena.capacity())
        return fail(PT_ERROR_INVALID_ARGUMENT, "the request's pixels exceed the server's memory budget");
    {
        std::scoped_lock lock(S.mu);
        if (S.stop) return fail(PT_ERROR_INVALID_STATE, "the server is stopping");
        if (!S.error.empty()) return fail(PT_ERROR_DEVICE, "the server failed: " + S.error);
        r->id = S.next_id++;
        r->t_submit = S.now_ms();
        S.pending.push_back(r);
        S.all[r->id] = r;
        S.submitted++;
    }
    S.cv_work.notify_one();
    id = r->id;
    return PT_OK;
}

static void fill_info(const ServerRequest& r, pt_request_info& info) {
    info.id = r.id;
    info.state = r.state;
    info.scene = r.rec.scene;
    info.samples = r.total;
    info.submit_ms = r.t_submit;
    info.admit_ms = r.t_admit;
    info.done_ms = r.t_done;
}

pt_status Renderer::ServerWait(uint64_t id, uint32_t timeout_ms, pt_request_info& info) {
    if (!server_) return fail(PT_ERROR_INVALID_STATE, "the server is not running");
    ServerState& S = *server_;
    std::unique_lock lock(S.mu);
    auto it = S.all.find(id);
    if (it == S.all.end()) return fail(PT_ERROR_INVALID_ARGUMENT, "unknown request " + std::to_string(id));
    std::shared_ptr<ServerRequest> r = it->second;
    S.cv_done.wait_for(lock, std::chrono::milliseconds(timeout_ms),
                       [&]() { return r->state == PT_REQUEST_DONE || r->state == PT_REQUEST_FAILED; });
    fill_info(*r, info);
    if (r->state == PT_REQUEST_FAILED && !r->message.empty())
        return fail(PT_ERROR_DEVICE, "request " + std::to_string(id) + ": " + r->message);
    return PT_OK;
}

static void fill_stats(ServerState& S, pt_server_stats& st) {
    std::memset(&st, 0, sizeof(st));
    st.requests_submitted = S.submitted;
    st.requests_done = S.done;
    st.requests_failed = S.failed;
    st.samples = S.samples_done;
    st.extension_rays = S.ext_rays;
    st.shadow_rays = S.shadow_rays;
    st.iterations = (uint64_t)S.iterations_seen;
    st.wall_ms = S.now_ms();
    st.busy_ms = S.busy_ms + (S.busy_since >= 0 ? st.wall_ms - S.busy_since : 0.0);
}

pt_status Renderer::ServerStats(pt_server_stats& stats) {
    if (!server_) return fail(PT_ERROR_INVALID_STATE, "the server is not running");
    std::scoped_lock lock(server_->mu);
    fill_stats(*server_, stats);
    return PT_OK;
}

pt_status Renderer::ServerStop(pt_server_stats* stats) {
    if (!server_) return fail(PT_ERROR_INVALID_STATE, "the server is not running");
    ServerState& S = *server_;
    {
        std::scoped_lock lock(S.mu);
        S.stop = true;
    }
    S.cv_work.notify_all();
    if (S.thread.joinable()) S.thread.join();
    S.cv_rb.notify_all();
    if (S.readback.joinable()) S.readback.join();
    std::string error;
    {
        std::scoped_lock lock(S.mu);
        if (stats) fill_stats(S, *stats);
        error = S.error;
        for (hipEvent_t e : S.free_events) (void)hipEventDestroy(e);
        S.free_events.clear();
    }
    (void)hipStreamSynchronize(stream_);
    if (S.rb_stream) (void)hipStreamDestroy(S.rb_stream);
    server_.reset();
    is_rendering_.store(false);
    if (!error.empty()) return fail(PT_ERROR_DEVICE, "the server failed: " + error);
    return PT_OK;
}

// ---------------------------------------------------------------- server thread

void Renderer::server_loop() {
    ServerState& S = *server_;
    const pt_status s = server_step();
    std::vector<std::shared_ptr<ServerRequest>> failed;
    {
        std::scoped_lock lock(S.mu);
        if (s != PT_OK) {
            S.error = last_error_;
            for (auto& r : S.pending) failed.push_back(r);
            for (auto& r : S.active) failed.push_back(r);
            S.pending.clear();
            S.active.clear();
            for (auto& r : failed) {
                r->state = PT_REQUEST_FAILED;
                r->t_done = S.now_ms();
                S.failed++;
            }
        }
        if (S.busy_since >= 0) {
            S.busy_ms += S.now_ms() - S.busy_since;
            S.busy_since = -1;
        }
        S.loop_exited = true;
    }
    (void)hipStreamSynchronize(stream_);
    S.cv_done.notify_all();
    S.cv_rb.notify_all();
}

// Runs until stopped and drained; returns the first error.
pt_status Renderer::server_step() {
    ServerState& S = *server_;
    const PathState p = path_state();
    const int lag = 1;   // counters of the launch before the one in flight
    uint64_t k = 0;   // graph launches (two iterations each)

    auto write_header = [&]() -> pt_status {
        struct Header {
            uint64_t total_samples;
            uint32_t view_lo, view_hi;
        } h;
        h.total_samples = S.total_samples;
        h.view_lo = (uint32_t)(S.chunks.empty() ? S.view_hi : S.chunks.front().virt);
        h.view_hi = (uint32_t)S.view_hi;
        static_assert(offsetof(FrameParams, view_hi) + 4 == sizeof(Header), "header of FrameParams");
        PT_TRY_HIP(launch_write_words(stream_, d_fp_.as<void>(), &h, sizeof(h)), "server frame header");
        S.header_dirty = false;
        return PT_OK;
    };

    while (true) {
        // ---- admission; sleep while there is nothing to run
        std::vector<std::shared_ptr<ServerRequest>> admitted;
        {
            std::unique_lock lock(S.mu);
            while (true) {
                while (!S.pending.empty() && S.active.size() < S.cfg.max_requests) {
                    auto& r = S.pending.front();
                    uint32_t off = 0;
                    if (!S.arena.alloc(r->pixels, off)) break;
                    r->accum_offset = off;
                    r->rec.accum_offset = off;
                    r->state = PT_REQUEST_RUNNING;
                    r->t_admit = S.now_ms();
                    admitted.push_back(r);
                    S.active.push_back(r);
                    S.pending.pop_front();
                }
                if (!S.active.empty()) {
                    if (S.busy_since < 0) S.busy_since = S.now_ms();
                    break;
                }
                if (S.busy_since >= 0) {
                    S.busy_ms += S.now_ms() - S.busy_since;
                    S.busy_since = -1;
                }
                if (S.stop && S.pending.empty()) return PT_OK;
                S.cv_work.wait(lock);
            }
        }
        for (auto& r : admitted) PT_TRY_HIP(launch_fill_accum(stream_, p, r->accum_offset, r->accum_offset + r->pixels), "clear request pixels");

        // ---- schedule chunks into the window, fewest remaining samples first
        uint64_t floor = S.view_hi;
        for (auto& r : S.active)
            if (r->scheduled > 0) floor = std::min(floor, r->first_virtual);
        uint64_t queued = S.total_samples - std::min(S.total_samples, S.launched_seen);
        while (queued < S.cfg.window_samples) {
            ServerRequest* best = nullptr;
            for (auto& r : S.active)
                if (r->scheduled < r->total && (!best || r->total - r->scheduled < best->total - best->scheduled))
                    best = r.get();
            if (!best || S.view_hi - floor >= kViewRing) break;
            const uint64_t n = std::min(best->total - best->scheduled, S.cfg.chunk_samples);
            ViewRecord rec = best->rec;
            rec.sample_begin = S.total_samples;
            rec.sample_count = n;
            rec.sample_offset = best->scheduled;
            PT_TRY_HIP(launch_write_words(stream_, d_views_.as<ViewRecord>() + (S.view_hi & (kViewRing - 1)), &rec,
                                          sizeof(rec)), "view record");
            if (best->scheduled == 0) {
                best->first_virtual = S.view_hi;
                floor = std::min(floor, S.view_hi);
            }
            best->last_virtual = S.view_hi;
            best->scheduled += n;
            std::shared_ptr<ServerRequest> owner;
            for (auto& r : S.active)
                if (r.get() == best) owner = r;
            S.chunks.push_back(ServerChunk{S.view_hi, S.total_samples + n, owner});
            S.total_samples += n;
            S.view_hi++;
            queued += n;
            S.header_dirty = true;
        }
        if (S.header_dirty) PT_TRY(write_header());

        // ---- kGraphIterations bounce iterations.  Every path in flight and
        // every queued sample belongs to an active request with scheduled
        // samples (a request leaves the active set only once its last path has
        // ended; one with nothing scheduled has nothing on the device), so
        // while those requests view one scene the launch may use the one-scene
        // traversal kernels, which read the top of that scene's tree from LDS.
        LoopMode mode = server_mode_;
        {
            uint32_t scene = kInvalid;
            bool one = true;
            for (auto& r : S.active) {
                if (r->scheduled == 0) continue;
                if (scene == kInvalid) scene = r->rec.scene;
                one = one && r->rec.scene == scene;
            }
            if (one && scene != kInvalid) {
                mode.trace_multi = false;
                mode.scene = scene;
            }
        }
        hipGraphExec_t graph = server_graph_;
        if (graph && !mode.trace_multi) {
            if (server_scene_graphs_[mode.scene]) graph = server_scene_graphs_[mode.scene];
            else mode = server_mode_;
        }
        if (graph) {
            PT_TRY_HIP(hipGraphLaunch(graph, stream_), "graph launch");
        } else {
            for (int i = 0; i < kGraphIterations; i++) PT_TRY(launch_iteration(i & 1, false, mode));
        }
        const int slot = (int)(k % kRing);
        PT_TRY_HIP(hipMemcpyAsync(h_ctr_ring_ + slot, p.ctr, sizeof(Counters), hipMemcpyDeviceToHost, stream_), "counters");
        PT_TRY_HIP(hipEventRecord(ring_events_[slot], stream_), "event");
        k++;
        if (k <= (uint64_t)lag) continue;

        // ---- the snapshot of launch k-1-lag
        const int old = (int)((k - 1 - lag) % kRing);
        PT_TRY_HIP(hipEventSynchronize(ring_events_[old]), "sync");
        const Counters c = h_ctr_ring_[old];
        const int64_t its = kGraphIterations * (int64_t)(k - lag);
        if (c.stack_overflow) return fail(PT_ERROR_DEVICE, "traversal stack overflow");
        {
            std::scoped_lock lock(S.mu);
            S.launched_seen = c.launched;
            S.iterations_seen = its;
            S.ext_rays = c.ext_rays;
            S.shadow_rays = c.shadow_rays;
        }
        while (!S.chunks.empty() && S.chunks.front().end <= c.launched) {
            ServerChunk& ch = S.chunks.front();
            if (ch.virt == ch.req->last_virtual && ch.req->scheduled == ch.req->total) ch.req->launched_at = its;
            S.chunks.pop_front();
            S.header_dirty = true;
        }
        // finished requests: tone-map in place, hand to the readback thread
        std::vector<std::shared_ptr<ServerRequest>> finished;
        for (auto& r : S.active)
            if (r->launched_at >= 0 && its >= r->launched_at + (int64_t)r->depth) finished.push_back(r);
        // requests seen finished in one snapshot go out in the order they finished
        std::stable_sort(finished.begin(), finished.end(), [](const auto& a, const auto& b) {
            return a->launched_at + (int64_t)a->depth < b->launched_at + (int64_t)b->depth;
        });
        for (auto& r : finished) {
            {
                std::unique_lock lock(S.mu);
                S.cv_work.wait(lock, [&]() { return !S.free_slots.empty(); });
                r->invalid_slot = S.free_slots.back();
                S.free_slots.pop_back();
            }
            // count this request's invalid pixels in its own counter, copied to
            // pinned memory before the event the readback waits on
            uint32_t* slot = d_req_invalid_.as<uint32_t>() + r->invalid_slot;
            PT_TRY_HIP(launch_zero_u32(stream_, slot, 1), "zero");
            PT_TRY_HIP(launch_postprocess(stream_, p, (uint32_t)r->last_virtual, (uint32_t)r->last_virtual + 1u,
                                          r->accum_offset, r->accum_offset + r->pixels, true, slot), "postprocess");
            PT_TRY_HIP(hipMemcpyAsync(h_req_invalid_ + r->invalid_slot, slot, sizeof(uint32_t), hipMemcpyDeviceToHost,
                                      stream_), "request counter");
            hipEvent_t e = nullptr;
            {
                std::scoped_lock lock(S.mu);
                if (!S.free_events.empty()) {
                    e = S.free_events.back();
                    S.free_events.pop_back();
                }
            }
            if (!e) PT_TRY_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming), "event");
            PT_TRY_HIP(hipEventRecord(e, stream_), "event");
            std::scoped_lock lock(S.mu);
            S.active.erase(std::find(S.active.begin(), S.active.end(), r));
            S.rb_queue.emplace_back(r, e);
        }
        if (!finished.empty()) S.cv_rb.notify_one();
    }
}

// ---------------------------------------------------------------- readback thread

void Renderer::readback_loop() {
    ServerState& S = *server_;
    while (true) {
        std::shared_ptr<ServerRequest> r;
        hipEvent_t e = nullptr;
        {
            std::unique_lock lock(S.mu);
            S.cv_rb.wait(lock, [&]() { return !S.rb_queue.empty() || S.loop_exited; });
            if (S.rb_queue.empty()) return;
            r = S.rb_queue.front().first;
            e = S.rb_queue.front().second;
            S.rb_queue.pop_front();
        }
        hipError_t err = hipEventSynchronize(e);
        const size_t np = r->pixels;
        if (err == hipSuccess && r->rgba)
            err = hipMemcpyAsync(r->rgba, d_frame_.as<uint32_t>() + r->accum_offset, np * sizeof(uint32_t),
                                 hipMemcpyDeviceToHost, S.rb_stream);
        if (err == hipSuccess && r->hdr)
            err = hipMemcpyAsync(r->hdr, d_accum_.as<float4>() + r->accum_offset, np * sizeof(float4),
                                 hipMemcpyDeviceToHost, S.rb_stream);
        if (err == hipSuccess) err = hipStreamSynchronize(S.rb_stream);
        const uint32_t invalid = h_req_invalid_[r->invalid_slot];
        {
            std::scoped_lock lock(S.mu);
            S.arena.release(r->accum_offset, r->pixels);
            S.free_events.push_back(e);
            S.free_slots.push_back(r->invalid_slot);
            r->t_done = S.now_ms();
            if (err == hipSuccess && invalid) {
                // RenderFrame's check of the same counter (Renderer::run_views)
                r->state = PT_REQUEST_FAILED;
                r->message = std::to_string(invalid) + " pixels hold non-finite or negative radiance";
                S.failed++;
            } else if (err == hipSuccess) {
                r->state = PT_REQUEST_DONE;
                S.done++;
                S.samples_done += r->total;
            } else {
                r->state = PT_REQUEST_FAILED;
                S.failed++;
                if (S.error.empty()) S.error = std::string("result readback: ") + hipGetErrorString(err);
            }
        }
        S.cv_done.notify_all();
        S.cv_work.notify_one();
    }
}

}  // namespace pt
