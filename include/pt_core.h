/*
 * pt_core.h - C ABI of the MI355X wavefront path-tracing core.
 *
 * This is the drop-in boundary for the reference renderer's scene-upload and
 * render entry points (Core/Include/Core/Graphics/Cuda/PathTracing/Renderer.hpp:114-143)
 * and, one level down, for its kernel launchers
 * (Core/Include/Core/Graphics/Cuda/PathTracing/Kernels/Launchers.hpp:18-146).
 *
 * Conventions
 *  - Every call that can fail returns a pt_status; PT_OK is 0.  The message of
 *    the last failed call on a handle is returned by pt_get_error_message().
 *    This replaces std::expected<void, Core::Utils::Error> (Core/Include/Core/Utils/Error.hpp:11-123).
 *  - Errors raised on the render thread are parked and read with
 *    pt_peek_last_error()/pt_consume_last_error(), as Renderer::PeekLastError /
 *    ConsumeLastError do (Renderer.cpp:308-320).
 *  - Plain pointers and sizes only.  Host pointers unless a name ends in
 *    _device, in which case the pointer is device memory of the renderer's GPU
 *    and `stream` is a hipStream_t (NULL = the renderer's own stream).
 *  - Calls are not thread-safe, except pt_frame_copy*, pt_get_progress,
 *    pt_is_rendering and the error accessors, as in the reference
 *    (Renderer.hpp:95-99,131-143).
 */
#ifndef PT_CORE_H
#define PT_CORE_H

#include <stddef.h>
#include <stdint.h>

/* The library is built with hidden visibility: only these entry points are exported. */
#if defined(__GNUC__) || defined(__clang__)
#pragma GCC visibility push(default)
#endif

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pt_renderer pt_renderer;

typedef enum pt_status {
    PT_OK = 0,
    PT_ERROR_INVALID_ARGUMENT = 1,
    PT_ERROR_INVALID_STATE = 2,   /* e.g. "Renderer is already rendering" (Renderer.cpp:347-348) */
    PT_ERROR_OUT_OF_RANGE = 3,    /* e.g. ResumeSimulation start frame (Renderer.cpp:406-407) */
    PT_ERROR_DEVICE = 4,          /* a HIP runtime call failed */
    PT_ERROR_OUT_OF_MEMORY = 5,
    PT_ERROR_UNSUPPORTED = 6,     /* shading model outside this core (Normal, Phong) */
    PT_ERROR_IO = 7               /* a file could not be read or written */
} pt_status;

/* Values of GlobalShadingModel (Core/Include/Core/Graphics/Cuda/PathTracing/Material.hpp:8-17). */
enum {
    PT_MODEL_NORMAL = 0,
    PT_MODEL_DIFFUSE = 1,
    PT_MODEL_MIRROR = 2,
    PT_MODEL_PHONG = 3,
    PT_MODEL_GGX = 4,
    PT_MODEL_EMISSIVE = 5,
    PT_MODEL_COUNT = 6
};

/* Material queues of the wavefront, in the order the shade launch visits them,
 * plus the regeneration queue for misses (MaterialEvalQueueViewsProvider.hpp:9-24,
 * RegenQueueView.hpp:10-85). */
enum {
    PT_QUEUE_DIFFUSE = 0,
    PT_QUEUE_GGX = 1,
    PT_QUEUE_MIRROR = 2,
    PT_QUEUE_EMISSIVE = 3,
    PT_QUEUE_REGEN = 4,
    PT_QUEUE_COUNT = 5
};

#define PT_INVALID_INDEX 0xFFFFFFFFu

/* Constant (untextured) material.  Replaces the reference's texture-handle
 * Material (Material.hpp:73-87): every channel is the fp32 value the 1x1
 * texture would return. */
typedef struct pt_material {
    uint32_t model;        /* PT_MODEL_* */
    float color[3];        /* albedo / base color, linear */
    float specular[3];     /* mirror reflectance (the reference default is 0) */
    float roughness;       /* rma.x */
    float metallic;        /* rma.y */
    float ior;             /* 1.0 default */
    float emission[3];     /* radiance */
    float normal_ts[3];    /* constant tangent-space normal, (2*texel-1); (0,0,1) = flat */
} pt_material;

/* Flattened world-space triangle soup.  Replaces the (entt::registry,
 * Assets::Storage) pair of InitializeSceneBuffers (Renderer.cpp:292-306):
 * the caller has already applied BuildTriangles' transforms (Renderer.cpp:159-213). */
typedef struct pt_scene_desc {
    uint32_t triangle_count;
    const float* positions;       /* [T][3 corners][xyz] */
    const float* normals;         /* [T][3][xyz], world space, unit length */
    const float* tangents;        /* [T][3][xyzw], w = handedness sign; NULL = derive from normals */
    const uint32_t* material_ids; /* [T] index into materials */
    uint32_t material_count;
    const pt_material* materials;
    const float* light_weights;   /* [T] alias-table weight of each emissive triangle;
                                     NULL = triangle area, as LightSampler.cpp:16-18 */
} pt_scene_desc;

/* Settings of StartSimulation (Renderer.hpp:118-126) and PathTracerDefaults.hpp:9-35. */
typedef struct pt_render_settings {
    uint32_t width;
    uint32_t height;
    uint32_t samples_per_pixel;  /* rounded down to a square, Renderer.cpp:366-367 */
    uint32_t path_depth;         /* clamped to [1,100] */
    uint32_t use_nee;            /* next-event estimation on/off */
    uint32_t rr_start_depth;     /* 0 = default 3 */
    uint64_t seed;               /* PCG32 seed of the sample streams */
} pt_render_settings;

/* Ecs::Camera (Core/Include/Core/Graphics/Ecs/Camera.hpp): vertical fov in radians. */
typedef struct pt_camera {
    float fov_y;
    float near_plane;
    float far_plane;
} pt_camera;

/* Capture::MotionState (Core/Include/Core/Capture/Sample.hpp:6-10). */
typedef struct pt_motion_state {
    float position[3];
    float rotation[4];   /* unit quaternion, x y z w */
} pt_motion_state;

typedef struct pt_progress {
    uint32_t is_rendering;
    uint32_t done_frames;
    uint32_t total_frames;
    uint32_t width;
    uint32_t height;
    uint32_t samples_per_pixel;
    uint32_t samples_per_pixel_axis;
    uint32_t path_depth;
    uint64_t done_samples;
    uint64_t total_samples;
} pt_progress;

/* Device-side counters of one rendered frame. */
typedef struct pt_frame_stats {
    double render_ms;            /* hipEvent time of the render loop (no upload, no readback) */
    uint64_t extension_rays;     /* rays fed to closest-hit traversal */
    uint64_t shadow_rays;        /* rays fed to any-hit traversal */
    uint64_t samples;            /* camera samples generated */
    uint32_t iterations;         /* wavefront iterations */
    uint32_t closest_launches;   /* closest-hit launches that had work */
    double closest_ms;           /* summed hipEvent time of closest-hit launches (profiled frames only) */
    double shadow_ms;            /* summed time of any-hit launches (profiled frames only) */
    double shade_ms;             /* summed time of shade launches (profiled frames only) */
    double regen_ms;             /* summed time of regeneration launches (profiled frames only) */
    uint64_t closest_nodes;      /* BVH nodes visited by closest-hit (instrumented frames only) */
    uint64_t closest_tris;       /* triangles tested by closest-hit (instrumented frames only) */
    uint64_t shadow_nodes;
    uint64_t shadow_tris;
    double bin_ms;               /* summed time of the queue-binning launches (profiled frames only) */
} pt_frame_stats;

/* One camera view of a resident scene: a frame of RenderFrame (Renderer.cpp:468-647)
 * as one element of a batch.  scene 0 is the scene of pt_initialize_scene_buffers;
 * pt_add_scene returns the others. */
typedef struct pt_view {
    uint32_t scene;
    uint32_t reserved;
    pt_render_settings settings;   /* width, height, spp, depth, NEE, RR start, seed */
    pt_camera camera;
    pt_motion_state state;
} pt_view;

/* Frame-mode flags for pt_render_frame and pt_render_batch. */
enum {
    PT_FRAME_DEFAULT = 0,
    PT_FRAME_PROFILE = 1,       /* time each stage with hipEvents (no graph replay) */
    PT_FRAME_INSTRUMENT = 2,    /* count nodes/triangles visited (instrumented traversal) */
    PT_FRAME_NO_GRAPH = 4       /* launch stage by stage instead of replaying the hipGraph */
};

/* Test/bench ray record: 32 bytes. */
typedef struct pt_ray {
    float origin[3];
    float t_min;
    float direction[3];
    float t_max;
} pt_ray;

/* Closest-hit record: original (pre-BVH) primitive index, PT_INVALID_INDEX on a miss. */
typedef struct pt_hit {
    uint32_t prim;
    float t;
    float u;
    float v;
} pt_hit;

/* Alias-table bin (AliasTableView.hpp:8-14): q, p, alias, value. */
typedef struct pt_alias_bin {
    float q;
    float p;
    uint32_t alias;
    uint32_t value;
} pt_alias_bin;

typedef struct pt_bvh_info {
    uint32_t node_count;      /* 128-byte BVH4 nodes */
    uint32_t leaf_count;
    uint32_t max_depth;       /* BVH4 levels */
    uint32_t max_leaf_size;
    uint32_t triangle_count;
    uint32_t light_count;
    uint32_t stack_lds;       /* traversal-stack entries held in LDS per lane */
    uint32_t stack_total;     /* worst-case stack entries (LDS + spill) */
    uint32_t node_width;      /* 4 or 2 children per node */
    uint32_t bvh8_node_count; /* 128-byte 8-wide quantized nodes traversed by any hit (0: none built) */
    uint32_t bvh8_depth;      /* their levels */
} pt_bvh_info;

/* Kernel launch and build parameters picked by the autotuner. */
typedef struct pt_tuning {
    uint32_t trace_block;        /* threads per block of the traversal kernels: 64, 128, 256 or 512 */
    uint32_t trace_blocks_per_cu;/* persistent grid of the closest-hit traversal (and of the any-hit one
                                    when shadow_blocks_per_cu is 0), blocks per CU: 0 = as many as fit */
    uint32_t shade_block;        /* 256 */
    uint32_t pool_size;          /* path slots (queue capacity): 1024..2^27, or 0 = automatic (default): the next
                                    power of two of the largest run so far, 2^22..2^27 slots (128 B each + queues) */
    uint32_t use_graph;          /* replay the bounce loop as a hipGraph */
    uint32_t bvh_max_leaf;       /* 1..16 triangles; applied at the next scene upload */
    float bvh_traversal_cost;    /* SAH cost of a node visit relative to a triangle test; next upload */
    uint32_t trace_refill;       /* idle lanes of a wave that trigger fetching new rays (1..64) */
    uint32_t bvh_width;          /* children per BVH node: 4 (128-byte nodes) or 2 (64-byte); next upload */
    uint32_t bvh_builder;        /* PT_BVH_BUILDER_*: where the next uploads build their BVH */
    uint32_t bvh_quantized;      /* traversals on 64-byte nodes whose child boxes are 8-bit offsets
                                    (conservative, results unchanged): bit 0 closest hit, bit 1 any
                                    hit; bit 2: any hit on 128-byte 8-wide quantized nodes instead,
                                    a collapse of the host builder's tree; bit 3 (with bit 2):
                                    closest hit on them as well; 4-wide trees only; next upload;
                                    default 14 */
    uint32_t shadow_blocks_per_cu;/* persistent grid of the any-hit traversal, blocks per CU; 0 =
                                    trace_blocks_per_cu if that is set, else as many as fit, less one
                                    wave per SIMD where the launch overlaps a closest-hit launch: in
                                    the replayed bounce loop each bounce's any-hit launch runs beside
                                    the next closest-hit launch, and the free slots let the two share
                                    the CUs */
    uint32_t stack_limit;        /* 0: each traversal's stack (LDS + global spill) holds the deepest
                                    resident BVH's worst case, so it cannot overflow; n > 0 caps it
                                    at n entries per ray.  A traversal whose stack would overflow
                                    drops the subtree and the call fails with PT_ERROR_DEVICE
                                    "traversal stack overflow" (a test of that path) */
} pt_tuning;

/* BVH builders: the host's binned-SAH build (best trees), or a linear BVH over
 * Morton codes built on the GPU (fast for scenes of millions of triangles; 4-wide
 * nodes only; leaves of at most min(bvh_max_leaf, 2) triangles).  Traversal
 * results do not depend on the tree. */
enum { PT_BVH_BUILDER_HOST = 0, PT_BVH_BUILDER_GPU = 1 };

/* ---- lifetime (Renderer ctor/dtor, Renderer.hpp:106-112) ---- */
pt_status pt_renderer_create(pt_renderer** out);
void pt_renderer_destroy(pt_renderer* renderer);
const char* pt_get_error_message(const pt_renderer* renderer);
const char* pt_version(void);

/* ---- scene upload and rendering buffers ---- */
/* Renderer::InitializeRenderingBuffers, Renderer.cpp:265-290 */
pt_status pt_initialize_rendering_buffers(pt_renderer* renderer, uint32_t width, uint32_t height);
/* Renderer::InitializeSceneBuffers, Renderer.cpp:292-306 */
pt_status pt_initialize_scene_buffers(pt_renderer* renderer, const pt_scene_desc* scene);
/* Renderer::Clear, Renderer.cpp:322-328 */
pt_status pt_clear(pt_renderer* renderer, uint8_t r, uint8_t g, uint8_t b, uint8_t a);

/* ---- asynchronous rendering (render thread) ---- */
/* Renderer::StartSimulation, Renderer.cpp:330-399.  output_folder may be NULL (no PNGs). */
pt_status pt_start_simulation(pt_renderer* renderer, const pt_render_settings* settings,
                              const pt_camera* camera, const pt_motion_state* motion_states,
                              uint32_t motion_state_count, const char* output_folder);
/* Renderer::ResumeSimulation, Renderer.cpp:401-427 */
pt_status pt_resume_simulation(pt_renderer* renderer, uint32_t start_frame);
/* Renderer::StopRendering, Renderer.cpp:429-440 */
pt_status pt_stop_rendering(pt_renderer* renderer);
/* Renderer::IsRendering and the progress getters, Renderer.hpp:131-140 */
int pt_is_rendering(const pt_renderer* renderer);
pt_status pt_get_progress(const pt_renderer* renderer, pt_progress* out);
/* FrameView::Lock()->GetData(), Renderer.hpp:45-104: RGBA8 pixels, row 0 = bottom. */
pt_status pt_frame_copy(pt_renderer* renderer, uint32_t* rgba, size_t pixel_count);
/* FrameView::Lock() without a copy (Renderer.hpp:45-104, LockedFrameView::GetData
 * and GetSize): takes the renderer's frame lock and hands out the current RGBA8
 * frame (row 0 = bottom; NULL and 0 before a frame size is set) until
 * pt_frame_unlock, which the same thread must call.  While it is held, a render
 * that finishes waits to publish its frame; on the locking thread every call that
 * takes the frame (pt_frame_copy, pt_clear, pt_save_frame, the render and
 * buffer-initialising calls) or waits for the render thread (pt_stop_rendering)
 * fails with PT_ERROR_INVALID_STATE instead of deadlocking.  pt_renderer_destroy
 * on the locking thread releases the lock first. */
pt_status pt_frame_lock(pt_renderer* renderer, const uint32_t** rgba, size_t* pixel_count);
pt_status pt_frame_unlock(pt_renderer* renderer);
/* accumulation / spp as float RGBA (needed for relMSE; no reference counterpart). */
pt_status pt_frame_copy_hdr(pt_renderer* renderer, float* rgba, size_t pixel_count);
/* The frame as an image file, by extension (SaveImage, ImageIO.cpp:119-160, as the
 * render loop's PNG save uses it, Renderer.cpp:649-675): ".png" writes the RGBA8
 * frame of pt_frame_copy, ".exr" the float radiance of pt_frame_copy_hdr (needs a
 * finished pt_render_frame); top row first.  Empty path, no extension ->
 * PT_ERROR_INVALID_ARGUMENT; another extension -> PT_ERROR_UNSUPPORTED. */
pt_status pt_save_frame(pt_renderer* renderer, const char* path);
/* Renderer::PeekLastError / ConsumeLastError, Renderer.cpp:308-320.
 * Copy the parked render-thread error into buf; returns its length, 0 if none. */
size_t pt_peek_last_error(const pt_renderer* renderer, char* buf, size_t capacity);
size_t pt_consume_last_error(pt_renderer* renderer, char* buf, size_t capacity);

/* ---- synchronous frame (bench and tests): one RenderFrame, Renderer.cpp:468-647 ---- */
pt_status pt_render_frame(pt_renderer* renderer, const pt_render_settings* settings,
                          const pt_camera* camera, const pt_motion_state* motion_state,
                          uint32_t frame_flags, pt_frame_stats* stats);

/* ---- resident scenes and batched views (no reference counterpart: the reference
 * renders one scene and one view per RenderFrame, Renderer.cpp:468-647) ---- */
/* Uploads a further scene (same validation as pt_initialize_scene_buffers) and
 * returns its id in 1..63. */
pt_status pt_add_scene(pt_renderer* renderer, const pt_scene_desc* scene, uint32_t* scene_id);
/* The frame size a render of `settings` produces (Renderer.cpp:354-372 clamps it):
 * the pixel count every result buffer of that view must hold. */
void pt_frame_size(const pt_render_settings* settings, uint32_t* width, uint32_t* height);
/* Frees a resident scene (id 0 included). */
pt_status pt_remove_scene(pt_renderer* renderer, uint32_t scene_id);
/* Renders view_count (1..4096) views, of any resident scenes, in one wavefront: the
 * bounce-loop kernels take paths of every view from one path pool.  Each view's
 * image equals the image it gets rendered alone.  hdr[i] (w*h float RGBA,
 * radiance / spp) and rgba8[i] (w*h RGBA8, row 0 = bottom) receive view i; either
 * array, or any entry, may be NULL.  w and h are the view's settings clamped as
 * RenderFrame clamps them (1..7680 x 1..4320, pt_frame_size).  stats aggregate
 * the whole batch. */
pt_status pt_render_batch(pt_renderer* renderer, const pt_view* views, uint32_t view_count,
                          float* const* hdr, uint32_t* const* rgba8, uint32_t frame_flags,
                          pt_frame_stats* stats);

/* ---- serving (no reference counterpart: the reference serves one request at a time
 * through StartSimulation/RenderFrame, Renderer.cpp:330-399,468-647) ----
 * A server thread keeps every resident scene on the device and runs one continuous
 * wavefront.  Submitted views are admitted while their pixels fit the memory budget;
 * admitted requests share the path pool and the material and shadow queues, their
 * samples scheduled in chunks, fewest remaining samples first.  A request's result
 * equals the view rendered alone.  Scenes cannot change while the server runs; the
 * other render calls return PT_ERROR_INVALID_STATE meanwhile. */
typedef struct pt_server_config {
    uint64_t accum_budget_bytes;  /* device memory for admitted requests' pixels, 20 B per
                                     pixel; 0 = 2 GiB.  A request larger than it is rejected. */
    uint64_t window_samples;      /* samples queued on the device ahead of the launch front;
                                     0 = four times the path pool */
    uint64_t chunk_samples;       /* largest piece of one request scheduled at once; 0 = pool/4 */
    uint32_t max_requests;        /* requests admitted at once, 1..2048; 0 = 1024 */
    uint32_t reserved;
} pt_server_config;

enum { PT_REQUEST_QUEUED = 0, PT_REQUEST_RUNNING = 1, PT_REQUEST_DONE = 2, PT_REQUEST_FAILED = 3 };

typedef struct pt_request_info {
    uint64_t id;
    uint32_t state;               /* PT_REQUEST_* */
    uint32_t scene;
    uint64_t samples;
    double submit_ms;             /* server clock: ms since pt_server_start */
    double admit_ms;              /* pixels allocated, first chunk may be scheduled */
    double done_ms;               /* results copied to the caller's buffers */
} pt_request_info;

typedef struct pt_server_stats {
    uint64_t requests_submitted;
    uint64_t requests_done;
    uint64_t requests_failed;
    uint64_t samples;             /* camera samples of finished requests */
    uint64_t extension_rays;      /* rays traced by the server's wavefront so far */
    uint64_t shadow_rays;
    uint64_t iterations;          /* bounce iterations run */
    double wall_ms;               /* server start to now (or to stop) */
    double busy_ms;               /* time with at least one admitted request */
} pt_server_stats;

pt_status pt_server_start(pt_renderer* renderer, const pt_server_config* config);
/* Queues one view; hdr (w*h float RGBA, radiance / spp) and rgba8 (w*h RGBA8, row 0 =
 * bottom) may be NULL and must stay valid until the request is done. */
pt_status pt_server_submit(pt_renderer* renderer, const pt_view* view, float* hdr, uint32_t* rgba8,
                           uint64_t* request_id);
/* Waits up to timeout_ms for the request to finish; info->state tells whether it did. */
pt_status pt_server_wait(pt_renderer* renderer, uint64_t request_id, uint32_t timeout_ms,
                         pt_request_info* info);
pt_status pt_server_get_stats(pt_renderer* renderer, pt_server_stats* stats);
/* Finishes every submitted request, then stops the server thread. */
pt_status pt_server_stop(pt_renderer* renderer, pt_server_stats* stats);

/* ---- learned path guiding (no reference counterpart: the reference samples
 * directions from the BSDF only, MaterialEval.cu:270-309,720-836) ----
 * An MLP of 3 hidden layers x 64 (BF16 weights, FP32 accumulation) maps a shading
 * point's position (with a sin/cos frequency encoding), normal and outgoing
 * direction to a mixture of 8 von Mises-Fisher lobes.  Diffuse and GGX vertices
 * sample it by one-sample MIS with the BSDF, so the estimator stays unbiased, and
 * the network is trained on the GPU with Adam from the radiance the paths carry. */
enum { PT_GUIDE_INPUTS = 64, PT_GUIDE_OUTPUTS = 64, PT_GUIDE_PARAMS = 16640 };

/* Training record of one scatter vertex (96 bytes). */
typedef struct pt_guide_record {
    float position[3];
    uint32_t prev;            /* previous record of the same path (0xFFFFFFFF: none) */
    float normal[3];          /* geometric normal facing the outgoing direction */
    float pdf;                /* density the direction was sampled with */
    float wo[3];              /* outgoing direction (toward the previous vertex) */
    float cos_i;              /* |shading normal . wi| */
    float wi[3];              /* sampled incident direction */
    float pad0;
    float throughput[3];      /* path throughput after this vertex */
    float pad1;
    float radiance[3];        /* incident radiance gathered along wi */
    float pad2;
} pt_guide_record;

typedef struct pt_guide_settings {
    uint32_t passes;          /* render passes (at most samples_per_pixel): pass k takes about 2^k / (2^passes - 1)
                                 of the samples per pixel (an even split when there are fewer than 2^passes - 1) */
    uint32_t train_steps;     /* Adam steps after each pass but the last */
    uint32_t batch;           /* records per Adam step (multiple of 64) */
    float learning_rate;
    float guide_fraction;     /* probability of sampling the guide; 0 = plain BSDF sampling */
    float weight_max;         /* clamp of a record's loss weight luminance(L) cos / pdf */
    uint32_t record_capacity; /* training records kept per pass, at most 2^26; 0 = 2^24.  Records come from an
                                 evenly spread subset of the samples (one in 2^k) sized to fit */
    uint32_t reset;           /* re-initialise the network before the first pass */
    uint64_t seed;            /* network initialisation and batch selection */
} pt_guide_settings;

typedef struct pt_guide_stats {
    double render_ms;         /* device time of the passes */
    double train_ms;          /* device time of the training steps */
    uint64_t records;         /* training records written by the last trained pass */
    float loss_first;         /* batch loss of the first and the last training step */
    float loss_last;
    uint32_t steps;           /* Adam steps taken in this call */
    uint32_t passes;
    double train_kernel_us;   /* average duration of one training forward+backward launch (the first 8 of each
                                 pass, each between its own events; the Adam step not included) */
    uint64_t shaded_entries;  /* diffuse and GGX entries shaded in the guided passes: the entries the guide
                                 network is evaluated for (a wave with any of them infers all 64 lanes) */
} pt_guide_stats;

/* Renders one frame in `passes` passes of growing size; the network is trained
 * between passes, and a freshly reset network is not sampled before its first
 * training.  The frame is read with pt_frame_copy and
 * pt_frame_copy_hdr; stats->render_ms includes the training time. */
pt_status pt_render_guided(pt_renderer* renderer, const pt_render_settings* settings,
                           const pt_camera* camera, const pt_motion_state* motion_state,
                           const pt_guide_settings* guide, pt_guide_stats* guide_stats,
                           pt_frame_stats* stats);
/* Network parameters, fp32: per layer W[64 out][64 in] then b[64]. */
pt_status pt_guide_reset(pt_renderer* renderer, uint64_t seed);
pt_status pt_guide_get_weights(pt_renderer* renderer, float* weights, uint32_t count);
pt_status pt_guide_set_weights(pt_renderer* renderer, const float* weights, uint32_t count);
/* Box that positions are normalised to (default: the bounds of scene 0). */
pt_status pt_guide_set_box(pt_renderer* renderer, const float lo[3], const float hi[3]);
/* Raw network outputs [n][64] of inputs [n][64] (the MFMA forward pass). */
pt_status pt_guide_forward(pt_renderer* renderer, const float* inputs, uint32_t count, float* outputs);
/* The mixtures the renderer's inference kernel stores for inputs [n][64], as it
 * stores them (fp16, decoded): lobes[n][8][5] = mu xyz, kappa, pi. */
pt_status pt_guide_mixture(pt_renderer* renderer, const float* inputs, uint32_t count, float* lobes);
/* Gradient (PT_GUIDE_PARAMS floats) of sum_i -w_i log q(wi_i) over the records, and
 * the loss, without changing the network. */
pt_status pt_guide_gradient(pt_renderer* renderer, const pt_guide_record* records, uint32_t count,
                            float weight_max, float* gradient, float* loss);
/* `steps` Adam steps on batches drawn from the records; losses[steps] may be NULL. */
pt_status pt_guide_train(pt_renderer* renderer, const pt_guide_record* records, uint32_t count,
                         uint32_t steps, uint32_t batch, float learning_rate, float weight_max,
                         float* losses);

/* ---- kernel-level hooks on caller-supplied batches ---- */
/* IntersectRaysWithSceneKernel (Intersect.cu:63-123) on a ray batch. */
pt_status pt_trace_closest(pt_renderer* renderer, const pt_ray* rays, uint32_t count, pt_hit* hits);
/* IntersectShadowRaysWithSceneKernel's occlusion test (Intersect.cu:136-179). occluded[i] = 0/1. */
pt_status pt_trace_any(pt_renderer* renderer, const pt_ray* rays, uint32_t count, uint32_t* occluded);
pt_status pt_trace_closest_device(pt_renderer* renderer, const pt_ray* rays_device, uint32_t count,
                                  pt_hit* hits_device, void* stream);
pt_status pt_trace_any_device(pt_renderer* renderer, const pt_ray* rays_device, uint32_t count,
                              uint32_t* occluded_device, void* stream);
/* pt_trace_closest_device that allocates nothing.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  Its
 * scratch (work counter, overflow flag, traversal stack spill) is one static
 * 32 MiB device array per GPU, held until the process ends; the persistent
 * grid shrinks until the deepest resident BVH's stack spill fits it, and a tree
 * whose spill does not fit one block is rejected with PT_ERROR_OUT_OF_RANGE.
 * Every call on one GPU shares that array: calls that overlap on one GPU race.
 * Asynchronous on `stream` (NULL: the renderer's stream); same hits as
 * pt_trace_closest_device. */
pt_status pt_trace_closest_device_noalloc(pt_renderer* renderer, const pt_ray* rays_device, uint32_t count,
                                          pt_hit* hits_device, void* stream);
/* Per-ray visit counts of the closest-hit traversal (the algorithmic-byte model). */
pt_status pt_trace_closest_counts(pt_renderer* renderer, const pt_ray* rays, uint32_t count,
                                  pt_hit* hits, uint32_t* nodes_visited, uint32_t* tris_tested);
/* The same for the any-hit traversal (the shadow rays' node kind, quantized by
 * default): occlusion and per-ray node and triangle visits. */
pt_status pt_trace_any_counts(pt_renderer* renderer, const pt_ray* rays, uint32_t count,
                              uint32_t* occluded, uint32_t* nodes_visited, uint32_t* tris_tested);
/* Queue membership of the closest-hit push (Intersect.cu:115-122): queue_of_ray[i] is a
 * PT_QUEUE_* value; queue_counts[PT_QUEUE_COUNT] receives the queue sizes. */
pt_status pt_classify_hits(pt_renderer* renderer, const pt_ray* rays, uint32_t count,
                           uint32_t* queue_of_ray, uint32_t* queue_counts);
/* AliasTableView::Sample (AliasTableView.hpp:32-60) of the light table on the GPU. */
pt_status pt_alias_sample(pt_renderer* renderer, const float* u, uint32_t count,
                          uint32_t* index, float* pmf);
/* Host copy of the light alias table built at scene upload (AliasTable.hpp:22-84). */
pt_status pt_get_alias_table(pt_renderer* renderer, pt_alias_bin* bins, uint32_t capacity,
                             uint32_t* count);
pt_status pt_get_bvh_info(pt_renderer* renderer, pt_bvh_info* info);

/* Queue occupancy of one wavefront iteration: what the reference reads back
 * every iteration to size the next one (Renderer.cpp:587-600), here kept for
 * each iteration of the last frame rendered with PT_FRAME_PROFILE. */
typedef struct pt_queue_occupancy {
    uint32_t extension;          /* rays traced by the iteration's closest-hit launch */
    uint32_t material[4];        /* entries of the diffuse, GGX, mirror and emissive queues */
    uint32_t missed;             /* rays that left the scene (their paths end) */
    uint32_t shadow;             /* shadow rays pushed by the shade */
    uint32_t next;               /* rays queued for the next iteration: continuing plus new camera paths */
    uint32_t capacity;           /* path slots, the capacity of every queue */
} pt_queue_occupancy;
pt_status pt_get_queue_occupancy(pt_renderer* renderer, pt_queue_occupancy* iterations, uint32_t capacity,
                                 uint32_t* count);
/* Raw BVH4 nodes (128 B each) and triangle order, for external traversal checks. */
pt_status pt_get_bvh4(pt_renderer* renderer, void* nodes, uint32_t node_capacity,
                      uint32_t* prim_order, uint32_t prim_capacity);

/* The quantized nodes of scene 0 (pt_tuning.bvh_quantized != 0 at upload), 64 B each,
 * as uploaded (node_count of pt_get_bvh_info); PT_ERROR_INVALID_STATE without them. */
pt_status pt_get_bvh4_quantized(pt_renderer* renderer, void* nodes, uint32_t node_capacity);
/* The 8-wide quantized nodes of scene 0 (pt_tuning.bvh_quantized bit 2 at a host-built
 * upload), 128 B each, as uploaded; *node_count receives their number (nodes == NULL:
 * only the count); prim_order (NULL or triangle-count entries) receives the original
 * prim of each position of the tree's triangle order.  PT_ERROR_INVALID_STATE without
 * them. */
pt_status pt_get_bvh8_quantized(pt_renderer* renderer, void* nodes, uint32_t node_capacity, uint32_t* node_count,
                                uint32_t* prim_order);

/* Host-only BVH build of the core (no GPU needed): the 128-byte BVH4 nodes and
 * BVH-order -> original prim permutation that InitializeSceneBuffers uploads.
 * Replaces HostBvh + DeviceBvh::BuildFlattenSync (HostBvh.cpp:51-100, DeviceBvh.cpp:7-52).
 * With nodes == NULL only *node_count and *depth are written. */
pt_status pt_build_bvh4(const float* positions, uint32_t triangle_count, uint32_t max_leaf_size,
                        void* nodes, uint32_t node_capacity, uint32_t* node_count,
                        uint32_t* prim_order, uint32_t* depth);

/* The same build collapsed into 8-wide quantized nodes (128 B each: box corner,
 * three scale exponents and the inner-child mask, 8-bit bounds of eight children,
 * the first inner child and first triangle, the children's triangle ranges; the
 * nodes the any-hit traversal walks when pt_tuning.bvh_quantized has bit 2),
 * numbered breadth first, each node's inner children consecutive and its leaf
 * children's triangles consecutive; prim_order receives that triangle order (the
 * same leaves as pt_build_bvh4, regrouped).  Host only; no reference counterpart.
 * With nodes == NULL only *node_count and *depth are written. */
pt_status pt_build_bvh8_quantized(const float* positions, uint32_t triangle_count, uint32_t max_leaf_size,
                                  void* nodes, uint32_t node_capacity, uint32_t* node_count,
                                  uint32_t* prim_order, uint32_t* depth);

/* The 64-byte quantized form of `count` BVH4 nodes of pt_build_bvh4 (the nodes
 * traversed when pt_tuning.bvh_quantized is set; no reference counterpart):
 * per node the box corner (3 floats), the three scale exponents, 8-bit child
 * bounds and the same child references.  Host only. */
pt_status pt_quantize_bvh4(const void* nodes, uint32_t count, void* quantized);

/* ---- scene ingestion (ObjLoader.cpp:67-234, Manager.cpp:311-412) ----
 * Reads a Wavefront OBJ and the MTL libraries it names (next to it) into a
 * scene for pt_initialize_scene_buffers / pt_add_scene, as the reference's
 * import does: faces fan-triangulated, per-vertex normals and uvs required,
 * triangles grouped per shape (o/g) and material; the MTL key `surface`
 * selects the model (unlit, unknown or absent: diffuse); Kd and Ks pass
 * through sRGB u8 quantisation (0.73 -> 222 -> 0.7304); roughness and metallic
 * are the default RMA texel (1, 0); Ke is kept as fp32; Ni is the ior; the
 * normal-map constant is (128, 128, 255); faces without a material get the
 * default material (white).  Tangents per submesh as pt_generate_tangents
 * computes them.  On failure *out is NULL and `error` receives the message. */
typedef struct pt_imported_scene pt_imported_scene;
pt_status pt_import_obj(const char* path, pt_imported_scene** out, char* error, size_t error_capacity);
/* The scene's description; valid until pt_imported_scene_free. */
const pt_scene_desc* pt_imported_scene_desc(const pt_imported_scene* scene);
void pt_imported_scene_free(pt_imported_scene* scene);
/* Per-vertex tangents of an indexed triangle mesh by Mikkelsen's method
 * (mikktspace, 180-degree threshold), as GenerateTangentSpace does at import
 * (TangentSpace.cpp:57-89): tangents[4 * v] = (x, y, z, sign); where faces
 * sharing a vertex disagree, the last face in index order sets it; vertices no
 * face uses get zeros.  positions/normals: float3 per vertex (normals unit),
 * uvs: float2 per vertex.  PT_ERROR_INVALID_ARGUMENT for an index count that
 * is not a multiple of 3 or an index >= vertex_count. */
pt_status pt_generate_tangents(const float* positions, const float* normals, const float* uvs,
                               uint32_t vertex_count, const uint32_t* indices, uint32_t index_count,
                               float* tangents);

/* ---- image files (ImageIO.cpp:119-160) ----
 * Pixels to a file whose extension picks the container: ".png" takes
 * PT_PIXELS_RGBA8 (little-endian RGBA8 words; an 8-bit RGBA PNG), ".exr" takes
 * PT_PIXELS_RGBA32F (float RGBA; an OpenEXR scanline file with FLOAT channels,
 * no compression).  bottom_up != 0: row 0 of `pixels` is the bottom row (the
 * frame buffers' order) and is written last.  Errors as SaveImage: empty path,
 * missing extension, zero width or height -> PT_ERROR_INVALID_ARGUMENT; another
 * extension, or pixels the container does not take -> PT_ERROR_UNSUPPORTED;
 * the file cannot be written -> PT_ERROR_IO.  No GPU is needed. */
typedef enum pt_pixel_format {
    PT_PIXELS_RGBA8 = 0,
    PT_PIXELS_RGBA32F = 1
} pt_pixel_format;
pt_status pt_write_image(const char* path, uint32_t width, uint32_t height, uint32_t pixel_format,
                         const void* pixels, int bottom_up);

/* ---- tuning ---- */
pt_status pt_get_tuning(pt_renderer* renderer, pt_tuning* tuning);
pt_status pt_set_tuning(pt_renderer* renderer, const pt_tuning* tuning);

#ifdef __cplusplus
}
#endif

#if defined(__GNUC__) || defined(__clang__)
#pragma GCC visibility pop
#endif

#endif /* PT_CORE_H */
