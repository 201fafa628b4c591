// Host side of learned path guiding: network state, test hooks and the
// multi-pass guided render with online training between passes.
#include <cmath>
#include <cstring>
#include <random>

#include "renderer.hpp"

namespace pt {

static_assert(sizeof(pt_guide_record) == sizeof(GuideRecord), "pt_guide_record mirrors GuideRecord");
static_assert(PT_GUIDE_PARAMS == kGuideParams, "parameter count");
static_assert(PT_GUIDE_INPUTS == kGuideDim && PT_GUIDE_OUTPUTS == kGuideDim, "layer width");

namespace {
constexpr uint64_t kMaxRecordCapacity = 1ull << 26;
constexpr uint64_t kDefaultRecordCapacity = 1ull << 24;
constexpr uint64_t kRecordsPerSample = 4;    // expected scatter vertices of a recording sample

// Samples per pixel of each pass: doubling from pass to pass (the later, better
// guided passes carry most of the image), or an even split when there are too
// few samples; each pass is a list of chunks with a square sample count (the
// stratified pixel grid of Renderer.cpp:366-367), smallest first.
std::vector<std::vector<uint32_t>> pass_chunks(uint32_t spp, uint32_t passes) {
    std::vector<uint32_t> n(passes);
    const uint64_t parts = passes < 32 ? (1ull << passes) - 1ull : ~0ull;
    uint64_t used = 0;
    if ((uint64_t)spp >= parts) {
        for (uint32_t k = 0; k + 1 < passes; k++) {
            n[k] = (uint32_t)((uint64_t)spp * (1ull << k) / parts);
            used += n[k];
        }
    } else {
        for (uint32_t k = 0; k + 1 < passes; k++) {
            n[k] = spp / passes;
            used += n[k];
        }
    }
    n[passes - 1] = (uint32_t)(spp - used);
    std::vector<std::vector<uint32_t>> out(passes);
    for (uint32_t k = 0; k < passes; k++) {
        uint32_t left = n[k];
        while (left > 0) {
            uint32_t a = (uint32_t)std::sqrt((double)left);
            while ((uint64_t)(a + 1) * (a + 1) <= left) a++;
            while ((uint64_t)a * a > left) a--;
            out[k].insert(out[k].begin(), a * a);
            left -= a * a;
        }
    }
    return out;
}
}  // namespace

// Guide box from the bounds of scene 0 (the union of the root's child boxes),
// unless the caller set one.
void Renderer::guide_box_from_scene() {
    if (g_box_set_ || !has_scene()) return;
    const Bvh4Node& root = host0().bvh.nodes[0];
    float lo[3] = {1e30f, 1e30f, 1e30f}, hi[3] = {-1e30f, -1e30f, -1e30f};
    for (int c = 0; c < 4; c++) {
        if (root.child[c] == kInvalid) continue;
        lo[0] = std::min(lo[0], root.lox[c]);
        lo[1] = std::min(lo[1], root.loy[c]);
        lo[2] = std::min(lo[2], root.loz[c]);
        hi[0] = std::max(hi[0], root.hix[c]);
        hi[1] = std::max(hi[1], root.hiy[c]);
        hi[2] = std::max(hi[2], root.hiz[c]);
    }
    GuideConfig c = g_cfg_;
    for (int a = 0; a < 3; a++) {
        if (!(hi[a] > lo[a])) {
            lo[a] = 0.0f;
            hi[a] = 1.0f;
        }
        c.lo[a] = lo[a];
        c.inv_extent[a] = 1.0f / (hi[a] - lo[a]);
    }
    if (std::memcmp(&c, &g_cfg_, sizeof(c)) != 0) {
        g_cfg_ = c;
        invalidate_graph();
    }
}

GuideDev Renderer::guide_dev() const {
    GuideDev g{};
    g.w32 = g_w32_.as<float>();
    g.w = g_w_.as<uint16_t>();
    g.wt = g_wt_.as<uint16_t>();
    g.grad = g_grad_.as<float>();
    g.m = g_m_.as<float>();
    g.v = g_v_.as<float>();
    g.wf = g_wf_.as<uint16_t>();
    g.bf = g_bf_.as<float>();
    g.recs = g_recs_.as<GuideRecord>();
    g.rec_capacity = g_rec_capacity_;
    g.head[0] = g_head_[0].as<uint32_t>();
    g.head[1] = g_head_[1].as<uint32_t>();
    g.sh_head = g_sh_head_.as<uint32_t>();
    g.rr_scale[0] = g_rr_scale_[0].as<float>();
    g.rr_scale[1] = g_rr_scale_[1].as<float>();
    g.cfg = g_cfg_;
    return g;
}

pt_status Renderer::ensure_guide() {
    if (g_w32_.bytes() == 0) {
        PT_TRY_HIP(g_w32_.allocate(sizeof(float) * kGuideParams), "guide weights");
        PT_TRY_HIP(g_w_.allocate(sizeof(uint16_t) * kGuideLayers * kGuideDim * kGuideDim), "guide weights");
        PT_TRY_HIP(g_wt_.allocate(sizeof(uint16_t) * kGuideLayers * kGuideDim * kGuideDim), "guide weights");
        PT_TRY_HIP(g_wf_.allocate(sizeof(uint16_t) * kGuideLayers * kGuideDim * kGuideDim), "guide weights");
        PT_TRY_HIP(g_bf_.allocate(sizeof(float) * kGuideLayers * kGuideDim), "guide biases");
        PT_TRY_HIP(g_grad_.allocate(sizeof(float) * kGuideParams), "guide gradient");
        PT_TRY_HIP(g_m_.allocate(sizeof(float) * kGuideParams), "guide Adam state");
        PT_TRY_HIP(g_v_.allocate(sizeof(float) * kGuideParams), "guide Adam state");
        if (!g_box_set_) {
            for (int a = 0; a < 3; a++) {
                g_cfg_.lo[a] = 0.0f;
                g_cfg_.inv_extent[a] = 1.0f;
            }
        }
    }
    guide_box_from_scene();
    if (!g_init_) PT_TRY(GuideReset(1));
    return PT_OK;
}

pt_status Renderer::ensure_guide_pool(uint32_t cap) {
    if (g_pool_ != pool_capacity_ || g_head_[0].bytes() == 0) {
        invalidate_graph();
        g_head_[0].release();
        g_head_[1].release();
        g_sh_head_.release();
        g_rr_scale_[0].release();
        g_rr_scale_[1].release();
        PT_TRY_HIP(g_head_[0].allocate(sizeof(uint32_t) * (size_t)pool_capacity_), "guide record heads");
        PT_TRY_HIP(g_head_[1].allocate(sizeof(uint32_t) * (size_t)pool_capacity_), "guide record heads");
        PT_TRY_HIP(g_sh_head_.allocate(sizeof(uint32_t) * (size_t)pool_capacity_), "guide record heads");
        PT_TRY_HIP(g_rr_scale_[0].allocate(sizeof(float) * (size_t)pool_capacity_), "guide roulette scales");
        PT_TRY_HIP(g_rr_scale_[1].allocate(sizeof(float) * (size_t)pool_capacity_), "guide roulette scales");
        g_pool_ = pool_capacity_;
    }
    if (cap != g_rec_capacity_ || g_recs_.bytes() == 0) {
        invalidate_graph();
        g_recs_.release();
        PT_TRY_HIP(g_recs_.allocate(sizeof(GuideRecord) * (size_t)cap), "guide records");
        g_rec_capacity_ = cap;
    }
    return PT_OK;
}

// He-uniform hidden layers (bound sqrt(6 / fan_in)), a 10x smaller output
// layer, zero biases except log kappa = 1 (broad initial lobes).
pt_status Renderer::GuideReset(uint64_t seed) {
    if (IsRendering()) return fail(PT_ERROR_INVALID_STATE, "Renderer is already rendering");
    g_init_ = true;   // ensure_guide below must not recurse into a reset
    PT_TRY(ensure_guide());
    std::vector<float> w(kGuideParams, 0.0f);
    std::mt19937_64 rng(seed);
    std::uniform_real_distribution<double> uni(-1.0, 1.0);
    for (int l = 0; l < kGuideLayers; l++) {
        const double fan_in = l == 0 ? (double)kGuideFeatures : (double)kGuideDim;
        double bound = std::sqrt(6.0 / fan_in);
        if (l == kGuideLayers - 1) bound *= 0.1;
        float* W = w.data() + (size_t)l * kGuideLayerParams;
        for (int o = 0; o < kGuideDim; o++)
            for (int i = 0; i < kGuideDim; i++)
                W[o * kGuideDim + i] = (l == 0 && i >= kGuideFeatures) ? 0.0f : (float)(bound * uni(rng));
        if (l == kGuideLayers - 1)
            for (int j = 0; j < kGuideLobes; j++) W[kGuideDim * kGuideDim + 5 * j + 3] = 1.0f;
    }
    return GuideSetWeights(w.data(), kGuideParams);
}

pt_status Renderer::GuideGetWeights(float* w, uint32_t n) {
    if (!w || n < (uint32_t)kGuideParams) return fail(PT_ERROR_INVALID_ARGUMENT, "weights buffer too small");
    PT_TRY(ensure_guide());
    PT_TRY_HIP(hipMemcpy(w, g_w32_.as<void>(), sizeof(float) * kGuideParams, hipMemcpyDeviceToHost), "guide weights");
    return PT_OK;
}

pt_status Renderer::GuideSetWeights(const float* w, uint32_t n) {
    if (IsRendering()) return fail(PT_ERROR_INVALID_STATE, "Renderer is already rendering");
    if (!w || n != (uint32_t)kGuideParams)
        return fail(PT_ERROR_INVALID_ARGUMENT, "expected " + std::to_string(kGuideParams) + " weights");
    for (uint32_t k = 0; k < n; k++)
        if (!std::isfinite(w[k])) return fail(PT_ERROR_INVALID_ARGUMENT, "non-finite weight");
    PT_TRY(ensure_guide());
    PT_TRY_HIP(hipMemcpyAsync(g_w32_.as<void>(), w, sizeof(float) * kGuideParams, hipMemcpyHostToDevice, stream_),
               "guide weights");
    PT_TRY_HIP(hipMemsetAsync(g_grad_.as<void>(), 0, sizeof(float) * kGuideParams, stream_), "guide gradient");
    PT_TRY_HIP(hipMemsetAsync(g_m_.as<void>(), 0, sizeof(float) * kGuideParams, stream_), "guide Adam state");
    PT_TRY_HIP(hipMemsetAsync(g_v_.as<void>(), 0, sizeof(float) * kGuideParams, stream_), "guide Adam state");
    PT_TRY_HIP(launch_guide_sync(stream_, guide_dev()), "guide weights");
    PT_TRY_HIP(hipStreamSynchronize(stream_), "guide weights");
    g_step_ = 0;
    g_init_ = true;
    return PT_OK;
}

pt_status Renderer::GuideSetBox(const float lo[3], const float hi[3]) {
    if (IsRendering()) return fail(PT_ERROR_INVALID_STATE, "Renderer is already rendering");
    for (int a = 0; a < 3; a++) {
        if (!(hi[a] > lo[a]) || !std::isfinite(lo[a]) || !std::isfinite(hi[a]))
            return fail(PT_ERROR_INVALID_ARGUMENT, "the guide box needs lo < hi on every axis");
        g_cfg_.lo[a] = lo[a];
        g_cfg_.inv_extent[a] = 1.0f / (hi[a] - lo[a]);
    }
    g_box_set_ = true;
    invalidate_graph();
    return PT_OK;
}

pt_status Renderer::GuideForward(const float* in, uint32_t n, float* out) {
    if (IsRendering()) return fail(PT_ERROR_INVALID_STATE, "Renderer is already rendering");
    if (n == 0) return PT_OK;
    if (!in || !out) return fail(PT_ERROR_INVALID_ARGUMENT, "NULL argument");
    PT_TRY(ensure_guide());
    const size_t bytes = (size_t)n * kGuideDim * sizeof(float);
    PT_TRY_HIP(d_hook_in_.allocate(bytes), "hook input");
    PT_TRY_HIP(d_hook_out_.allocate(bytes), "hook output");
    PT_TRY_HIP(hipMemcpyAsync(d_hook_in_.as<void>(), in, bytes, hipMemcpyHostToDevice, stream_), "hook input");
    PT_TRY_HIP(launch_guide_forward(stream_, guide_dev(), d_hook_in_.as<float>(), n, d_hook_out_.as<float>()), "guide forward");
    PT_TRY_HIP(hipMemcpyAsync(out, d_hook_out_.as<void>(), bytes, hipMemcpyDeviceToHost, stream_), "hook output");
    PT_TRY_HIP(hipStreamSynchronize(stream_), "hook sync");
    return PT_OK;
}

pt_status Renderer::GuideMixture(const float* in, uint32_t n, float* lobes) {
    if (IsRendering()) return fail(PT_ERROR_INVALID_STATE, "Renderer is already rendering");
    if (n == 0) return PT_OK;
    if (!in || !lobes) return fail(PT_ERROR_INVALID_ARGUMENT, "NULL argument");
    PT_TRY(ensure_guide());
    const size_t in_bytes = (size_t)n * kGuideDim * sizeof(float);
    const size_t words = (size_t)n * kGuideMixWords;
    PT_TRY_HIP(d_hook_in_.allocate(in_bytes), "hook input");
    PT_TRY_HIP(d_hook_out_.allocate(words * sizeof(uint32_t)), "hook output");
    PT_TRY_HIP(hipMemcpyAsync(d_hook_in_.as<void>(), in, in_bytes, hipMemcpyHostToDevice, stream_), "hook input");
    PT_TRY_HIP(launch_guide_mixture(stream_, guide_dev(), d_hook_in_.as<float>(), n, d_hook_out_.as<uint32_t>()),
               "guide mixture");
    std::vector<uint32_t> w(words);
    PT_TRY_HIP(hipMemcpyAsync(w.data(), d_hook_out_.as<void>(), words * sizeof(uint32_t), hipMemcpyDeviceToHost, stream_),
               "hook output");
    PT_TRY_HIP(hipStreamSynchronize(stream_), "hook sync");
    // word 4 c + t of entry e holds component c of lobes t (low half) and t + 4
    for (size_t e = 0; e < n; e++)
        for (int j = 0; j < kGuideLobes; j++)
            for (int c = 0; c < 5; c++) {
                const uint32_t v = w[(size_t)(4 * c + (j & 3)) * n + e];
                lobes[(e * kGuideLobes + j) * 5 + c] = j < 4 ? guide_half_lo(v) : guide_half_hi(v);
            }
    return PT_OK;
}

pt_status Renderer::GuideGradient(const pt_guide_record* recs, uint32_t n, float w_max, float* grad, float* loss) {
    if (IsRendering()) return fail(PT_ERROR_INVALID_STATE, "Renderer is already rendering");
    if (!recs || !grad || !loss) return fail(PT_ERROR_INVALID_ARGUMENT, "NULL argument");
    PT_TRY(ensure_guide());
    PT_TRY_HIP(g_loss_.allocate(sizeof(float) * 16), "guide loss");
    PT_TRY_HIP(d_hook_in_.allocate(sizeof(GuideRecord) * (size_t)std::max(n, 1u)), "hook input");
    PT_TRY_HIP(hipMemcpyAsync(d_hook_in_.as<void>(), recs, sizeof(GuideRecord) * (size_t)n, hipMemcpyHostToDevice, stream_), "hook input");
    PT_TRY_HIP(g_part_.allocate(sizeof(float) * kGuideParams * guide_parts(std::max(n, 1u))), "guide partial gradients");
    PT_TRY_HIP(hipMemsetAsync(g_loss_.as<void>(), 0, sizeof(float), stream_), "guide loss");
    PT_TRY_HIP(launch_guide_train(stream_, guide_dev(), d_hook_in_.as<GuideRecord>(), n, n, 0, w_max, g_part_.as<float>(),
                                  g_loss_.as<float>()),
               "guide gradient");
    if (n > 0) {
        PT_TRY_HIP(launch_guide_sum(stream_, g_part_.as<float>(), guide_parts(n), g_grad_.as<float>()), "guide gradient");
    } else {
        PT_TRY_HIP(hipMemsetAsync(g_grad_.as<void>(), 0, sizeof(float) * kGuideParams, stream_), "guide gradient");
    }
    PT_TRY_HIP(hipMemcpyAsync(grad, g_grad_.as<void>(), sizeof(float) * kGuideParams, hipMemcpyDeviceToHost, stream_), "guide gradient");
    PT_TRY_HIP(hipMemcpyAsync(loss, g_loss_.as<void>(), sizeof(float), hipMemcpyDeviceToHost, stream_), "guide loss");
    PT_TRY_HIP(hipStreamSynchronize(stream_), "hook sync");
    return PT_OK;
}

pt_status Renderer::GuideTrain(const pt_guide_record* recs, uint32_t n, uint32_t steps, uint32_t batch, float lr,
                               float w_max, float* losses) {
    if (IsRendering()) return fail(PT_ERROR_INVALID_STATE, "Renderer is already rendering");
    if (!recs || n == 0) return fail(PT_ERROR_INVALID_ARGUMENT, "no records");
    if (batch == 0 || batch % 64 != 0) return fail(PT_ERROR_INVALID_ARGUMENT, "batch must be a positive multiple of 64");
    if (!(lr > 0.0f)) return fail(PT_ERROR_INVALID_ARGUMENT, "learning rate must be positive");
    PT_TRY(ensure_guide());
    PT_TRY_HIP(g_loss_.allocate(sizeof(float) * std::max<size_t>(steps, 16)), "guide loss");
    PT_TRY_HIP(d_hook_in_.allocate(sizeof(GuideRecord) * (size_t)n), "hook input");
    PT_TRY_HIP(hipMemcpyAsync(d_hook_in_.as<void>(), recs, sizeof(GuideRecord) * (size_t)n, hipMemcpyHostToDevice, stream_), "hook input");
    PT_TRY_HIP(hipMemsetAsync(g_loss_.as<void>(), 0, sizeof(float) * std::max<size_t>(steps, 16), stream_), "guide loss");
    PT_TRY_HIP(g_part_.allocate(sizeof(float) * kGuideParams * guide_parts(batch)), "guide partial gradients");
    const GuideDev g = guide_dev();
    for (uint32_t s = 0; s < steps; s++) {
        g_step_++;
        const uint32_t seed = (g_step_ * 2654435761u) | 1u;
        PT_TRY_HIP(launch_guide_train(stream_, g, d_hook_in_.as<GuideRecord>(), n, batch, seed, w_max,
                                      g_part_.as<float>(), g_loss_.as<float>() + s),
                   "guide training");
        PT_TRY_HIP(launch_guide_adam(stream_, g, g_part_.as<float>(), guide_parts(batch), lr, g_step_, 1.0f / (float)batch),
                   "guide Adam step");
    }
    if (losses && steps) {
        PT_TRY_HIP(hipMemcpyAsync(losses, g_loss_.as<void>(), sizeof(float) * steps, hipMemcpyDeviceToHost, stream_), "guide loss");
    }
    PT_TRY_HIP(hipStreamSynchronize(stream_), "guide training");
    if (losses)
        for (uint32_t s = 0; s < steps; s++) losses[s] /= (float)batch;
    return PT_OK;
}

// ---------------------------------------------------------------- guided frame

pt_status Renderer::RenderGuided(const pt_render_settings& s, const pt_camera& camera, const pt_motion_state& state,
                                 const pt_guide_settings& gs, pt_guide_stats* gstats, pt_frame_stats* stats) {
    if (IsRendering()) return fail(PT_ERROR_INVALID_STATE, "Renderer is already rendering");
    if (!has_scene()) return fail(PT_ERROR_INVALID_STATE, "scene buffers are not initialized");
    if (gs.passes < 1) return fail(PT_ERROR_INVALID_ARGUMENT, "at least one pass");
    if (!(gs.guide_fraction >= 0.0f && gs.guide_fraction <= 1.0f))
        return fail(PT_ERROR_INVALID_ARGUMENT, "guide_fraction must be in [0, 1]");
    const bool learn = gs.guide_fraction > 0.0f;
    if (learn && gs.train_steps > 0 && (gs.batch == 0 || gs.batch % 64 != 0))
        return fail(PT_ERROR_INVALID_ARGUMENT, "batch must be a positive multiple of 64");
    if (learn && gs.train_steps > 0 && !(gs.learning_rate > 0.0f))
        return fail(PT_ERROR_INVALID_ARGUMENT, "learning rate must be positive");

    // the clamped (square) samples per pixel, spread over the passes' chunks
    pt_view v{};
    v.scene = 0;
    v.settings = s;
    v.camera = camera;
    v.state = state;
    std::vector<ViewRecord> views(1);
    uint32_t depth = 0;
    PT_TRY(make_view(v, views[0], depth));
    const uint32_t spp_total = views[0].spp;
    if (spp_total < gs.passes) return fail(PT_ERROR_INVALID_ARGUMENT, "fewer samples per pixel than passes");
    const std::vector<std::vector<uint32_t>> chunks = pass_chunks(spp_total, gs.passes);
    PT_TRY(set_frame_size(views[0].width, views[0].height));
    last_spp_ = spp_total;
    hdr_valid_ = false;
    const uint64_t pixels = (uint64_t)views[0].width * views[0].height;

    // the pool for the largest pass chunk, before the guide's per-slot arrays
    // are sized to it (the passes' own runs then find it large enough)
    uint32_t max_chunk = 0;
    for (const auto& pass : chunks)
        for (uint32_t c : pass) max_chunk = std::max(max_chunk, c);
    PT_TRY(ensure_pool((uint64_t)max_chunk * pixels));
    uint32_t rec_capacity = 0;
    if (learn) {
        guide_box_from_scene();
        PT_TRY(ensure_guide());
        if (gs.reset) PT_TRY(GuideReset(gs.seed));
        rec_capacity = gs.record_capacity ? (uint32_t)std::min<uint64_t>(gs.record_capacity, kMaxRecordCapacity)
                                          : (uint32_t)kDefaultRecordCapacity;
        PT_TRY(ensure_guide_pool(std::max<uint32_t>(rec_capacity, 1024)));
        PT_TRY_HIP(g_loss_.allocate(sizeof(float) * std::max<size_t>(gs.train_steps, 16)), "guide loss");
    }
    hipEvent_t t0 = nullptr, t1 = nullptr;
    PT_TRY_HIP(hipEventCreate(&t0), "event");
    PT_TRY_HIP(hipEventCreate(&t1), "event");
    // the first kTimedTrainSteps training launches of each pass are bracketed
    // by their own events: the training kernel's duration without the Adam step
    constexpr uint32_t kTimedTrainSteps = 8;
    hipEvent_t tk[2 * kTimedTrainSteps] = {};
    struct EventGuard {
        hipEvent_t a, b;
        hipEvent_t* k;
        ~EventGuard() {
            (void)hipEventDestroy(a);
            (void)hipEventDestroy(b);
            for (uint32_t i = 0; i < 2 * kTimedTrainSteps; i++)
                if (k[i]) (void)hipEventDestroy(k[i]);
        }
    } guard{t0, t1, tk};
    for (auto& e : tk) PT_TRY_HIP(hipEventCreate(&e), "event");
    double train_kernel_ms = 0.0;
    uint32_t train_kernel_n = 0;

    pt_guide_stats gst{};
    pt_frame_stats total{};
    bool trained = learn && g_step_ > 0 && !gs.reset;
    float first_loss = 0.0f, last_loss = 0.0f;
    uint32_t chunk_index = 0;
    for (uint32_t pass = 0; pass < gs.passes; pass++) {
        for (size_t ci = 0; ci < chunks[pass].size(); ci++, chunk_index++) {
            v.settings.samples_per_pixel = chunks[pass][ci];
            PT_TRY(make_view(v, views[0], depth));
            views[0].seed = s.seed + 0x9E3779B97F4A7C15ull * chunk_index;
            views[0].inv_spp = 1.0f / (float)spp_total;
            RunOptions opt;
            opt.clear = chunk_index == 0;
            opt.guided = learn;
            opt.guide_fraction = trained ? gs.guide_fraction : 0.0f;
            // training records from one sample in 2^rec_shift, so that the
            // expected vertices of the chunk fit the record buffer
            const uint64_t want = (uint64_t)chunks[pass][ci] * pixels * kRecordsPerSample;
            while (opt.rec_shift < 31 && (want >> opt.rec_shift) > rec_capacity) opt.rec_shift++;
            pt_frame_stats ps{};
            PT_TRY(run_views(views, depth, PT_FRAME_DEFAULT, &ps, nullptr, opt));
            total.render_ms += ps.render_ms;
            total.extension_rays += ps.extension_rays;
            total.shadow_rays += ps.shadow_rays;
            total.samples += ps.samples;
            total.iterations += ps.iterations;
            gst.render_ms += ps.render_ms;
            if (opt.guide_fraction > 0.0f) gst.shaded_entries += h_ctr_ring_[0].dg_entries;
        }
        if (!learn || pass + 1 == gs.passes || gs.train_steps == 0) continue;

        // train on the records of the pass's last (largest) chunk; run_views
        // left the run's final counters in h_ctr_ring_[0]
        const uint32_t count = std::min(h_ctr_ring_[0].records, g_rec_capacity_);
        gst.records = count;
        if (count == 0) continue;
        const GuideDev g = guide_dev();
        PT_TRY_HIP(g_part_.allocate(sizeof(float) * kGuideParams * guide_parts(gs.batch)), "guide partial gradients");
        PT_TRY_HIP(hipMemsetAsync(g_loss_.as<void>(), 0, sizeof(float) * gs.train_steps, stream_), "guide loss");
        PT_TRY_HIP(hipEventRecord(t0, stream_), "event");
        for (uint32_t k = 0; k < gs.train_steps; k++) {
            g_step_++;
            const uint32_t seed = (uint32_t)(gs.seed * 0x9E3779B1u) ^ (g_step_ * 2654435761u) ^ 1u;
            if (k < kTimedTrainSteps) PT_TRY_HIP(hipEventRecord(tk[2 * k], stream_), "event");
            PT_TRY_HIP(launch_guide_train(stream_, g, g.recs, count, gs.batch, seed | 1u,
                                          gs.weight_max > 0.0f ? gs.weight_max : 1e30f, g_part_.as<float>(),
                                          g_loss_.as<float>() + k),
                       "guide training");
            if (k < kTimedTrainSteps) PT_TRY_HIP(hipEventRecord(tk[2 * k + 1], stream_), "event");
            PT_TRY_HIP(launch_guide_adam(stream_, g, g_part_.as<float>(), guide_parts(gs.batch), gs.learning_rate, g_step_,
                                         1.0f / (float)gs.batch),
                       "guide Adam step");
        }
        PT_TRY_HIP(hipEventRecord(t1, stream_), "event");
        std::vector<float> losses(gs.train_steps);
        PT_TRY_HIP(hipMemcpyAsync(losses.data(), g_loss_.as<void>(), sizeof(float) * gs.train_steps, hipMemcpyDeviceToHost, stream_), "guide loss");
        PT_TRY_HIP(hipStreamSynchronize(stream_), "guide training");
        float ms = 0.0f;
        PT_TRY_HIP(hipEventElapsedTime(&ms, t0, t1), "elapsed");
        gst.train_ms += ms;
        for (uint32_t k = 0; k < std::min(gs.train_steps, kTimedTrainSteps); k++) {
            float km = 0.0f;
            PT_TRY_HIP(hipEventElapsedTime(&km, tk[2 * k], tk[2 * k + 1]), "elapsed");
            train_kernel_ms += km;
            train_kernel_n++;
        }
        total.render_ms += ms;
        if (gst.steps == 0) first_loss = losses[0] / (float)gs.batch;
        last_loss = losses[gs.train_steps - 1] / (float)gs.batch;
        gst.steps += gs.train_steps;
        trained = true;
    }
    gst.passes = gs.passes;
    gst.train_kernel_us = train_kernel_n ? 1e3 * train_kernel_ms / train_kernel_n : 0.0;
    gst.loss_first = first_loss;
    gst.loss_last = last_loss;

    PT_TRY(publish_frame());
    hdr_valid_ = true;
    if (stats) *stats = total;
    if (gstats) *gstats = gst;
    return PT_OK;
}

}  // namespace pt
