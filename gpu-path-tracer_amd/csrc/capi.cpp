// extern "C" entry points of include/pt_core.h: argument checks, exception
// barrier, and forwarding to pt::Renderer.
#include <cstring>
#include <new>

#include "image_io.hpp"
#include "renderer.hpp"

struct pt_renderer {
    pt::Renderer impl;
};

namespace {
template <class F>
pt_status guarded(pt_renderer* r, F&& f) {
    if (!r) return PT_ERROR_INVALID_ARGUMENT;
    try {
        return f(r->impl);
    } catch (const std::bad_alloc&) {
        return r->impl.fail(PT_ERROR_OUT_OF_MEMORY, "host allocation failed");
    } catch (const std::exception& e) {
        return r->impl.fail(PT_ERROR_INVALID_ARGUMENT, e.what());
    }
}
}  // namespace

extern "C" {

const char* pt_version(void) { return "pt-core 0.1 (gfx950 wavefront)"; }

pt_status pt_renderer_create(pt_renderer** out) {
    if (!out) return PT_ERROR_INVALID_ARGUMENT;
    *out = nullptr;
    pt_renderer* r = new (std::nothrow) pt_renderer();
    if (!r) return PT_ERROR_OUT_OF_MEMORY;
    const pt_status s = r->impl.init();
    if (s != PT_OK) {
        delete r;
        return s;
    }
    *out = r;
    return PT_OK;
}

void pt_renderer_destroy(pt_renderer* r) { delete r; }

const char* pt_get_error_message(const pt_renderer* r) { return r ? r->impl.error_message() : "null renderer"; }

pt_status pt_initialize_rendering_buffers(pt_renderer* r, uint32_t width, uint32_t height) {
    return guarded(r, [&](pt::Renderer& x) { return x.InitializeRenderingBuffers(width, height); });
}

pt_status pt_initialize_scene_buffers(pt_renderer* r, const pt_scene_desc* scene) {
    return guarded(r, [&](pt::Renderer& x) {
        if (!scene) return x.fail(PT_ERROR_INVALID_ARGUMENT, "scene is NULL");
        return x.InitializeSceneBuffers(*scene);
    });
}

pt_status pt_clear(pt_renderer* r, uint8_t cr, uint8_t cg, uint8_t cb, uint8_t ca) {
    return guarded(r, [&](pt::Renderer& x) { return x.Clear(cr, cg, cb, ca); });
}

pt_status pt_start_simulation(pt_renderer* r, const pt_render_settings* settings, const pt_camera* camera,
                              const pt_motion_state* motion_states, uint32_t motion_state_count,
                              const char* output_folder) {
    return guarded(r, [&](pt::Renderer& x) {
        if (!settings || !camera) return x.fail(PT_ERROR_INVALID_ARGUMENT, "settings and camera must not be NULL");
        return x.StartSimulation(*settings, *camera, motion_states, motion_state_count, output_folder);
    });
}

pt_status pt_resume_simulation(pt_renderer* r, uint32_t start_frame) {
    return guarded(r, [&](pt::Renderer& x) { return x.ResumeSimulation(start_frame); });
}

pt_status pt_stop_rendering(pt_renderer* r) {
    return guarded(r, [&](pt::Renderer& x) { return x.StopRendering(); });
}

int pt_is_rendering(const pt_renderer* r) { return r && r->impl.IsRendering() ? 1 : 0; }

pt_status pt_get_progress(const pt_renderer* r, pt_progress* out) {
    if (!r || !out) return PT_ERROR_INVALID_ARGUMENT;
    r->impl.GetProgress(*out);
    return PT_OK;
}

pt_status pt_frame_copy(pt_renderer* r, uint32_t* rgba, size_t pixel_count) {
    return guarded(r, [&](pt::Renderer& x) { return x.CopyFrame(rgba, pixel_count); });
}

pt_status pt_frame_lock(pt_renderer* r, const uint32_t** rgba, size_t* pixel_count) {
    return guarded(r, [&](pt::Renderer& x) { return x.LockFrame(rgba, pixel_count); });
}

pt_status pt_frame_unlock(pt_renderer* r) {
    return guarded(r, [&](pt::Renderer& x) { return x.UnlockFrame(); });
}

pt_status pt_frame_copy_hdr(pt_renderer* r, float* rgba, size_t pixel_count) {
    return guarded(r, [&](pt::Renderer& x) { return x.CopyFrameHdr(rgba, pixel_count); });
}

pt_status pt_get_bvh4_quantized(pt_renderer* r, void* nodes, uint32_t node_capacity) {
    return guarded(r, [&](pt::Renderer& x) { return x.GetBvh4Quantized(nodes, node_capacity); });
}

pt_status pt_quantize_bvh4(const void* nodes, uint32_t count, void* quantized) {
    if (count && (!nodes || !quantized)) return PT_ERROR_INVALID_ARGUMENT;
    const pt::Bvh4Node* in = static_cast<const pt::Bvh4Node*>(nodes);
    pt::Bvh4QNode* out = static_cast<pt::Bvh4QNode*>(quantized);
    for (uint32_t i = 0; i < count; i++) pt::quantize_bvh4_node(in[i], out[i]);
    return PT_OK;
}

pt_status pt_save_frame(pt_renderer* r, const char* path) {
    return guarded(r, [&](pt::Renderer& x) { return x.SaveFrame(path); });
}

static pt_status write_image(const char* path, uint32_t width, uint32_t height, uint32_t pixel_format,
                             const void* pixels, int bottom_up) {
    const std::string p = path ? path : "";
    pt::ImageKind kind;
    std::string err;
    pt::ImageStatus s = pt::image_kind(p, width, height, kind, err);
    if (s == pt::ImageStatus::Ok) {
        if (!pixels) return PT_ERROR_INVALID_ARGUMENT;
        if (kind == pt::ImageKind::Png && pixel_format == PT_PIXELS_RGBA8)
            s = pt::write_png_rgba8(p, width, height, (const uint32_t*)pixels, bottom_up != 0, err);
        else if (kind == pt::ImageKind::Exr && pixel_format == PT_PIXELS_RGBA32F)
            s = pt::write_exr_rgba32f(p, width, height, (const float*)pixels, bottom_up != 0, err);
        else
            return PT_ERROR_UNSUPPORTED;
    }
    switch (s) {
        case pt::ImageStatus::Ok: return PT_OK;
        case pt::ImageStatus::InvalidArgument: return PT_ERROR_INVALID_ARGUMENT;
        case pt::ImageStatus::Unsupported: return PT_ERROR_UNSUPPORTED;
        default: return PT_ERROR_IO;
    }
}

// ImageIO::SaveImage (ImageIO.cpp:119-160); no exception crosses the C ABI:
// the encoders build whole-image buffers (bad_alloc) and std::filesystem throws
pt_status pt_write_image(const char* path, uint32_t width, uint32_t height, uint32_t pixel_format,
                         const void* pixels, int bottom_up) {
    try {
        return write_image(path, width, height, pixel_format, pixels, bottom_up);
    } catch (const std::bad_alloc&) {
        return PT_ERROR_OUT_OF_MEMORY;
    } catch (...) {
        return PT_ERROR_IO;
    }
}

size_t pt_peek_last_error(const pt_renderer* r, char* buf, size_t capacity) {
    return r ? r->impl.PeekLastError(buf, capacity) : 0;
}

size_t pt_consume_last_error(pt_renderer* r, char* buf, size_t capacity) {
    return r ? r->impl.ConsumeLastError(buf, capacity) : 0;
}

pt_status pt_render_frame(pt_renderer* r, const pt_render_settings* settings, const pt_camera* camera,
                          const pt_motion_state* motion_state, uint32_t frame_flags, pt_frame_stats* stats) {
    return guarded(r, [&](pt::Renderer& x) {
        if (!settings || !camera || !motion_state) return x.fail(PT_ERROR_INVALID_ARGUMENT, "NULL argument");
        return x.RenderFrameSync(*settings, *camera, *motion_state, frame_flags, stats);
    });
}

pt_status pt_add_scene(pt_renderer* r, const pt_scene_desc* scene, uint32_t* scene_id) {
    return guarded(r, [&](pt::Renderer& x) {
        if (!scene || !scene_id) return x.fail(PT_ERROR_INVALID_ARGUMENT, "NULL argument");
        return x.AddScene(*scene, *scene_id);
    });
}

pt_status pt_remove_scene(pt_renderer* r, uint32_t scene_id) {
    return guarded(r, [&](pt::Renderer& x) { return x.RemoveScene(scene_id); });
}

pt_status pt_render_batch(pt_renderer* r, const pt_view* views, uint32_t view_count, float* const* hdr,
                          uint32_t* const* rgba8, uint32_t frame_flags, pt_frame_stats* stats) {
    return guarded(r, [&](pt::Renderer& x) { return x.RenderBatch(views, view_count, hdr, rgba8, frame_flags, stats); });
}

pt_status pt_server_start(pt_renderer* r, const pt_server_config* config) {
    return guarded(r, [&](pt::Renderer& x) {
        pt_server_config c{};
        if (config) c = *config;
        return x.ServerStart(c);
    });
}

pt_status pt_server_submit(pt_renderer* r, const pt_view* view, float* hdr, uint32_t* rgba8, uint64_t* request_id) {
    return guarded(r, [&](pt::Renderer& x) {
        if (!view || !request_id) return x.fail(PT_ERROR_INVALID_ARGUMENT, "NULL argument");
        return x.ServerSubmit(*view, hdr, rgba8, *request_id);
    });
}

pt_status pt_server_wait(pt_renderer* r, uint64_t request_id, uint32_t timeout_ms, pt_request_info* info) {
    return guarded(r, [&](pt::Renderer& x) {
        if (!info) return x.fail(PT_ERROR_INVALID_ARGUMENT, "NULL argument");
        return x.ServerWait(request_id, timeout_ms, *info);
    });
}

pt_status pt_server_get_stats(pt_renderer* r, pt_server_stats* stats) {
    return guarded(r, [&](pt::Renderer& x) {
        if (!stats) return x.fail(PT_ERROR_INVALID_ARGUMENT, "NULL argument");
        return x.ServerStats(*stats);
    });
}

pt_status pt_server_stop(pt_renderer* r, pt_server_stats* stats) {
    return guarded(r, [&](pt::Renderer& x) { return x.ServerStop(stats); });
}

pt_status pt_render_guided(pt_renderer* r, const pt_render_settings* settings, const pt_camera* camera,
                           const pt_motion_state* motion_state, const pt_guide_settings* guide,
                           pt_guide_stats* guide_stats, pt_frame_stats* stats) {
    return guarded(r, [&](pt::Renderer& x) {
        if (!settings || !camera || !motion_state || !guide) return x.fail(PT_ERROR_INVALID_ARGUMENT, "NULL argument");
        return x.RenderGuided(*settings, *camera, *motion_state, *guide, guide_stats, stats);
    });
}

pt_status pt_guide_reset(pt_renderer* r, uint64_t seed) {
    return guarded(r, [&](pt::Renderer& x) { return x.GuideReset(seed); });
}

pt_status pt_guide_get_weights(pt_renderer* r, float* weights, uint32_t count) {
    return guarded(r, [&](pt::Renderer& x) { return x.GuideGetWeights(weights, count); });
}

pt_status pt_guide_set_weights(pt_renderer* r, const float* weights, uint32_t count) {
    return guarded(r, [&](pt::Renderer& x) { return x.GuideSetWeights(weights, count); });
}

pt_status pt_guide_set_box(pt_renderer* r, const float lo[3], const float hi[3]) {
    return guarded(r, [&](pt::Renderer& x) {
        if (!lo || !hi) return x.fail(PT_ERROR_INVALID_ARGUMENT, "NULL argument");
        return x.GuideSetBox(lo, hi);
    });
}

pt_status pt_guide_forward(pt_renderer* r, const float* inputs, uint32_t count, float* outputs) {
    return guarded(r, [&](pt::Renderer& x) { return x.GuideForward(inputs, count, outputs); });
}

pt_status pt_guide_mixture(pt_renderer* r, const float* inputs, uint32_t count, float* lobes) {
    return guarded(r, [&](pt::Renderer& x) { return x.GuideMixture(inputs, count, lobes); });
}

pt_status pt_guide_gradient(pt_renderer* r, const pt_guide_record* records, uint32_t count, float weight_max,
                            float* gradient, float* loss) {
    return guarded(r, [&](pt::Renderer& x) { return x.GuideGradient(records, count, weight_max, gradient, loss); });
}

pt_status pt_guide_train(pt_renderer* r, const pt_guide_record* records, uint32_t count, uint32_t steps,
                         uint32_t batch, float learning_rate, float weight_max, float* losses) {
    return guarded(r, [&](pt::Renderer& x) {
        return x.GuideTrain(records, count, steps, batch, learning_rate, weight_max, losses);
    });
}

pt_status pt_trace_closest(pt_renderer* r, const pt_ray* rays, uint32_t count, pt_hit* hits) {
    return guarded(r, [&](pt::Renderer& x) { return x.TraceBatch(rays, count, false, false, hits, nullptr, nullptr, nullptr); });
}

pt_status pt_trace_any(pt_renderer* r, const pt_ray* rays, uint32_t count, uint32_t* occluded) {
    return guarded(r, [&](pt::Renderer& x) { return x.TraceBatch(rays, count, true, false, occluded, nullptr, nullptr, nullptr); });
}

pt_status pt_trace_closest_device(pt_renderer* r, const pt_ray* rays_device, uint32_t count, pt_hit* hits_device,
                                  void* stream) {
    return guarded(r, [&](pt::Renderer& x) {
        return x.TraceBatch(rays_device, count, false, true, hits_device, nullptr, nullptr, (hipStream_t)stream);
    });
}

pt_status pt_trace_closest_device_noalloc(pt_renderer* r, const pt_ray* rays_device, uint32_t count,
                                          pt_hit* hits_device, void* stream) {
    return guarded(r, [&](pt::Renderer& x) {
        return x.TraceClosestNoAlloc(rays_device, count, hits_device, (hipStream_t)stream);
    });
}

pt_status pt_trace_any_device(pt_renderer* r, const pt_ray* rays_device, uint32_t count, uint32_t* occluded_device,
                              void* stream) {
    return guarded(r, [&](pt::Renderer& x) {
        return x.TraceBatch(rays_device, count, true, true, occluded_device, nullptr, nullptr, (hipStream_t)stream);
    });
}

pt_status pt_trace_closest_counts(pt_renderer* r, const pt_ray* rays, uint32_t count, pt_hit* hits,
                                  uint32_t* nodes_visited, uint32_t* tris_tested) {
    return guarded(r, [&](pt::Renderer& x) {
        if (!nodes_visited || !tris_tested) return x.fail(PT_ERROR_INVALID_ARGUMENT, "count arrays must not be NULL");
        return x.TraceBatch(rays, count, false, false, hits, nodes_visited, tris_tested, nullptr);
    });
}

void pt_frame_size(const pt_render_settings* settings, uint32_t* width, uint32_t* height) {
    uint32_t w = 1, h = 1;
    if (settings) pt::frame_size(*settings, w, h);
    if (width) *width = w;
    if (height) *height = h;
}

pt_status pt_trace_any_counts(pt_renderer* r, const pt_ray* rays, uint32_t count, uint32_t* occluded,
                              uint32_t* nodes_visited, uint32_t* tris_tested) {
    return guarded(r, [&](pt::Renderer& x) {
        if (!nodes_visited || !tris_tested) return x.fail(PT_ERROR_INVALID_ARGUMENT, "count arrays must not be NULL");
        return x.TraceBatch(rays, count, true, false, occluded, nodes_visited, tris_tested, nullptr);
    });
}

pt_status pt_classify_hits(pt_renderer* r, const pt_ray* rays, uint32_t count, uint32_t* queue_of_ray,
                           uint32_t* queue_counts) {
    return guarded(r, [&](pt::Renderer& x) { return x.Classify(rays, count, queue_of_ray, queue_counts); });
}

pt_status pt_alias_sample(pt_renderer* r, const float* u, uint32_t count, uint32_t* index, float* pmf) {
    return guarded(r, [&](pt::Renderer& x) { return x.AliasSample(u, count, index, pmf); });
}

pt_status pt_get_alias_table(pt_renderer* r, pt_alias_bin* bins, uint32_t capacity, uint32_t* count) {
    return guarded(r, [&](pt::Renderer& x) { return x.GetAliasTable(bins, capacity, count); });
}

pt_status pt_get_bvh_info(pt_renderer* r, pt_bvh_info* info) {
    return guarded(r, [&](pt::Renderer& x) {
        if (!info) return x.fail(PT_ERROR_INVALID_ARGUMENT, "info is NULL");
        return x.GetBvhInfo(*info);
    });
}

pt_status pt_get_bvh8_quantized(pt_renderer* r, void* nodes, uint32_t node_capacity, uint32_t* node_count,
                                uint32_t* prim_order) {
    return guarded(r, [&](pt::Renderer& x) {
        if (!node_count) return x.fail(PT_ERROR_INVALID_ARGUMENT, "node_count is NULL");
        return x.GetBvh8Quantized(nodes, node_capacity, *node_count, prim_order);
    });
}

pt_status pt_get_queue_occupancy(pt_renderer* r, pt_queue_occupancy* iterations, uint32_t capacity, uint32_t* count) {
    return guarded(r, [&](pt::Renderer& x) {
        if (!count) return x.fail(PT_ERROR_INVALID_ARGUMENT, "count is NULL");
        return x.GetQueueOccupancy(iterations, capacity, *count);
    });
}

pt_status pt_get_bvh4(pt_renderer* r, void* nodes, uint32_t node_capacity, uint32_t* prim_order, uint32_t prim_capacity) {
    return guarded(r, [&](pt::Renderer& x) { return x.GetBvh4(nodes, node_capacity, prim_order, prim_capacity); });
}

pt_status pt_get_tuning(pt_renderer* r, pt_tuning* tuning) {
    return guarded(r, [&](pt::Renderer& x) {
        if (!tuning) return x.fail(PT_ERROR_INVALID_ARGUMENT, "tuning is NULL");
        return x.GetTuning(*tuning);
    });
}

pt_status pt_set_tuning(pt_renderer* r, const pt_tuning* tuning) {
    return guarded(r, [&](pt::Renderer& x) {
        if (!tuning) return x.fail(PT_ERROR_INVALID_ARGUMENT, "tuning is NULL");
        return x.SetTuning(*tuning);
    });
}

pt_status pt_build_bvh8_quantized(const float* positions, uint32_t triangle_count, uint32_t max_leaf_size,
                                  void* nodes, uint32_t node_capacity, uint32_t* node_count, uint32_t* prim_order,
                                  uint32_t* depth) {
    if ((triangle_count > 0 && !positions) || !node_count) return PT_ERROR_INVALID_ARGUMENT;
    try {
        pt::BvhBuildOptions o;
        if (max_leaf_size) o.max_leaf_size = max_leaf_size;
        o.wide8 = true;
        const pt::Bvh4Build b = pt::build_bvh4(positions, triangle_count, o);
        *node_count = (uint32_t)b.qnodes8.size();
        if (depth) *depth = b.depth8;
        if (!nodes) return PT_OK;
        if (node_capacity < b.qnodes8.size()) return PT_ERROR_INVALID_ARGUMENT;
        std::memcpy(nodes, b.qnodes8.data(), b.qnodes8.size() * sizeof(pt::Bvh8QNode));
        if (prim_order) std::memcpy(prim_order, b.order8.data(), b.order8.size() * sizeof(uint32_t));
        return PT_OK;
    } catch (const std::bad_alloc&) {
        return PT_ERROR_OUT_OF_MEMORY;
    } catch (const std::exception&) {
        return PT_ERROR_INVALID_ARGUMENT;
    }
}

pt_status pt_build_bvh4(const float* positions, uint32_t triangle_count, uint32_t max_leaf_size, void* nodes,
                        uint32_t node_capacity, uint32_t* node_count, uint32_t* prim_order, uint32_t* depth) {
    if ((triangle_count > 0 && !positions) || !node_count) return PT_ERROR_INVALID_ARGUMENT;
    try {
        pt::BvhBuildOptions o;
        if (max_leaf_size) o.max_leaf_size = max_leaf_size;
        const pt::Bvh4Build b = pt::build_bvh4(positions, triangle_count, o);
        *node_count = (uint32_t)b.nodes.size();
        if (depth) *depth = b.depth;
        if (!nodes) return PT_OK;
        if (node_capacity < b.nodes.size()) return PT_ERROR_INVALID_ARGUMENT;
        std::memcpy(nodes, b.nodes.data(), b.nodes.size() * sizeof(pt::Bvh4Node));
        if (prim_order) std::memcpy(prim_order, b.order.data(), b.order.size() * sizeof(uint32_t));
        return PT_OK;
    } catch (const std::exception&) {
        return PT_ERROR_INVALID_ARGUMENT;
    }
}

}  // extern "C"
