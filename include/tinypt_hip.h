/*
 * tinypt_hip.h — C ABI of the MI355X (gfx950) path-tracing backend, `device=hip`.
 *
 * This is the drop-in boundary that sits beside tinypt's `cpu` and `cuda`
 * backends.  Every entry point takes plain pointers and sizes; no C++ or torch
 * type crosses it.  The scene is handed over as a flat description
 * (`tpt_scene_desc`) built on the host from tinypt's object model; the library
 * builds its own acceleration structure, uploads one device arena and renders.
 *
 * Reference interfaces replaced (li-plus/tinypt):
 *   tpt_hip_scene_create / _destroy  <- tinypt::Scene::to(Device)           src/tinypt/scene.h:24-37
 *                                      cuda::Scene(const cpu::Scene&)       src/tinypt/cuda/scene.h:95-197
 *   tpt_hip_render                   <- tinypt::PathTracer::render()        src/tinypt/path_tracer.h:26-44
 *                                      cuda::PathTracer::render             src/tinypt/cuda/path_tracer.cpp:8-11
 *   tpt_hip_trace_closest            <- cpu::Scene::intersect (closest hit) src/tinypt/cpu/scene.h:27
 *   tpt_hip_trace_occluded           <- cpu::Scene::intersect_t (any hit)   src/tinypt/cpu/scene.h:28
 *   tpt_hip_last_error               <- CHECK_CUDA -> std::runtime_error    src/tinypt/cuda/defs.h:33-40
 *
 * Conventions shared with the reference CPU path (src/tinypt/cpu/):
 *   - z-up world, rotations stored as 3x3 with COLUMNS u, v, w (cpu/mapping.h:14-41);
 *   - images are row-major with row 0 at the BOTTOM (y up), as cpu::Image (cpu/image.h:13-94);
 *   - INF = 1e20, EPS = 1e-3 (cpu/defs.h:23-24).
 *
 * Thread-safety: one render per scene handle at a time; several handles may share a device.
 */
#ifndef TINYPT_HIP_H_
#define TINYPT_HIP_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TPT_ABI_VERSION 3

/* ---- error codes ------------------------------------------------------- */
enum {
    TPT_OK = 0,
    TPT_ERR_INVALID_ARGUMENT = -1, /* maps to std::invalid_argument */
    TPT_ERR_RUNTIME = -2,          /* maps to std::runtime_error    */
    TPT_ERR_HIP = -3,              /* HIP runtime failure (no device, OOM, launch) */
    TPT_ERR_UNSUPPORTED = -4       /* object / material / light kind not supported */
};

/* ---- textures ------------------------------------------------------------ */
enum { TPT_IMG_F32 = 0, TPT_IMG_U8 = 1 };

/* One texel image.  `channels` is 1 (alpha / height), 2 (bump gradient) or 3/4
 * (rgb / rgba).  U8 texels are read as value/255.  Row 0 is the bottom row. */
typedef struct {
    int32_t width, height, channels, format;
    const void *data;
} tpt_image;

/* RGBTexture / AlphaTexture (cpu/texture.h:12-63): a constant or a map.
 * For an alpha texture read from an RGBA image, `channel` selects the component. */
typedef struct {
    int32_t image;  /* -1: constant `value`; else index into tpt_scene_desc.images */
    int32_t channel;/* first channel read from the image (0 for rgb, 3 for rgba alpha) */
    float value[3]; /* rgb constant; an alpha constant uses value[0] */
} tpt_texture;

/* ---- materials (cpu/material.h:21-216) ------------------------------------ */
enum {
    TPT_BXDF_LAMBERTIAN = 0,
    TPT_BXDF_METAL = 1,
    TPT_BXDF_DIELECTRIC = 2,
    TPT_BXDF_GLOSSY_BLINN = 3, /* Glossy + BlinnPhongModel(shininess)        */
    TPT_BXDF_GLOSSY_GGX = 4    /* Glossy + TrowbridgeReitzModel(2/(s+2))     */
};
typedef struct {
    int32_t type;
    float param; /* dielectric: ior; glossy: shininess */
    tpt_texture color;
} tpt_bxdf;

enum { TPT_COMBINE_NONE = 0, TPT_COMBINE_ADD = 1, TPT_COMBINE_MIX = 2 };
typedef struct {
    tpt_bxdf lobe[2];     /* lobe[1] used only when combine != NONE (AddBXDF / MixBXDF) */
    int32_t combine;
    float weight;         /* MixBXDF weight of lobe[0] */
    tpt_texture emission; /* RGBTexture */
    tpt_texture alpha;    /* AlphaTexture; constant 1 = opaque */
    int32_t bump_image;   /* -1, or index of a 2-channel F32 gradient image (Sobel of the height map) */
} tpt_material;

/* ---- objects (cpu/object.h:22-291) ---------------------------------------- */
enum { TPT_OBJ_SPHERE = 0, TPT_OBJ_RECTANGLE = 1, TPT_OBJ_CIRCLE = 2, TPT_OBJ_MESH = 3 };
typedef struct {
    int32_t type;
    int32_t material; /* analytic objects: index into materials */
    int32_t mesh;     /* TPT_OBJ_MESH: index into meshes */
    int32_t pad_;
    float center[3];
    float radius;    /* sphere, circle */
    float rot[9];    /* rectangle, circle: columns u, v, w (w = normal) */
    float dim[2];    /* rectangle */
} tpt_object;

/* A triangle mesh, already in world space.  Each face holds 9 ints: for
 * corners a, b, c the (vertex, normal, texcoord) indices, -1 when absent
 * (tinyobj index_t, cpu/object.h:244-252).  Faces are grouped into `shapes`
 * (each shape owned one BVHTree(faces, 4) in the reference, cpu/object.cpp:590-609). */
typedef struct {
    int32_t num_vertices, num_normals, num_texcoords, num_faces, num_shapes;
    const float *vertices;        /* num_vertices * 3 */
    const float *normals;         /* num_normals * 3 */
    const float *texcoords;       /* num_texcoords * 2 */
    const int32_t *faces;         /* num_faces * 9 */
    const int32_t *face_material; /* num_faces, index into tpt_scene_desc.materials */
    const int32_t *shape_offsets; /* num_shapes + 1 face offsets */
} tpt_mesh;

/* ---- lights (cpu/light.h:12-143) ------------------------------------------ */
enum { TPT_LIGHT_DISTANT = 0, TPT_LIGHT_POINT = 1 };
typedef struct {
    int32_t type;
    float rot[9];    /* distant: columns u, v, w; w points towards the light */
    float center[3]; /* point */
    float color[3];
    float power;
    float radius;    /* point: sphere radius; distant: angle (full cone angle, radians) */
} tpt_delta_light;

/* Non-delta lights passed to Scene() (cpu/scene.cpp:21-27; ObjectLight and
 * EnvLight, cpu/light.h:30-43, 113-124).  Each joins the object-light mixture,
 * in order, after the scene's emissive objects and before the background's own
 * EnvLight.  TPT_OBJECT_LIGHT_OBJECT: ObjectLight of tpt_scene_desc.objects[object]
 * (a mesh samples a shape, then a face, uniformly: ObjectGroup::sample,
 * cpu/object.cpp:352-365); TPT_OBJECT_LIGHT_ENV: EnvLight of the scene's background. */
enum { TPT_OBJECT_LIGHT_OBJECT = 0, TPT_OBJECT_LIGHT_ENV = 1 };
typedef struct {
    int32_t kind;
    int32_t object;
} tpt_object_light;

/* ---- camera (cpu/camera.h:14-79) and background (cpu/texture.h:83-109) ---- */
typedef struct {
    float center[3];
    float rot[9];
    int32_t width, height;
    float fov, aperture, focus_dist;
} tpt_camera;

typedef struct {
    tpt_texture rgb;
    float rot[9];
    /* 0: reference-compatible EnvTexture (its luminance CDF is never built, so
     *    sampling is uniform in uv with pdf 1/(4 pi), texture.cpp:66-67);
     * 1: build the max-channel CDF the reference intended (texture.cpp:68-76). */
    int32_t importance;
} tpt_background;

typedef struct {
    tpt_camera camera;
    tpt_background background;
    int32_t num_objects, num_meshes, num_materials, num_images, num_delta_lights;
    const tpt_object *objects; /* top-level objects, in the order given to Scene() */
    const tpt_mesh *meshes;
    const tpt_material *materials;
    const tpt_image *images;
    const tpt_delta_light *delta_lights;
    int32_t num_object_lights;
    const tpt_object_light *object_lights; /* non-delta lights passed to Scene(), in order */
} tpt_scene_desc;

/* ---- render parameters ---------------------------------------------------- */
enum { TPT_OUT_LDR_CLIPPED = 0, TPT_OUT_HDR_MEAN = 1 };
typedef struct {
    int32_t spp;            /* samples per pixel requested (integer-divided over sub-pixels) */
    int32_t max_depth;      /* reference: 5 (cpu/path_tracer.cpp:103) */
    int32_t rr_depth;       /* < 0: no Russian roulette (reference); else RR once depth >= rr_depth */
    int32_t shadow_bounded; /* 0: reference unbounded delta-light shadow rays; 1: bounded by light distance */
    uint64_t seed;          /* base of the per-(pixel, sample) erand48 streams */
    int32_t output;         /* TPT_OUT_* */
    int32_t accum_fp64;     /* accumulate samples in FP64 (convergence references) */
    int32_t queue_capacity; /* max paths in flight; 0 = library default */
    int32_t block_size;     /* threads per block for the wavefront kernels; 0 = default (256) */
    int32_t count_traffic;  /* 1: count BVH node / triangle fetches in extend (slower; fills stats) */
} tpt_render_params;

typedef struct {
    double render_ms;       /* wall time of the render, host clock */
    double extend_ms;       /* summed device time of extend launches (HIP events) */
    double shade_ms, shadow_ms, other_ms;
    uint64_t paths;         /* paths traced = W * H * effective spp */
    uint64_t rays_closest;  /* closest-hit queries */
    uint64_t rays_shadow;   /* shadow (any-hit) queries */
    uint64_t extend_launches;
    uint64_t nodes_visited; /* BVH node fetches in extend, when counting is enabled */
    uint64_t tris_tested;   /* triangle records fetched in extend, when counting is enabled */
} tpt_render_stats;

/* ---- opaque handle -------------------------------------------------------- */
typedef struct tpt_hip_scene tpt_hip_scene;

/* Library ABI version (TPT_ABI_VERSION). */
int tpt_hip_abi_version(void);
/* Number of visible HIP devices (0 when none); never throws. */
int tpt_hip_device_count(void);
/* Message of the last failing call on this thread ("" if none). */
const char *tpt_hip_last_error(void);

/* Scene::to(hip): build the BVH on the host, upload one device arena on the
 * current device.  Replaces cuda::Scene(const cpu::Scene&) (cuda/scene.h:95-174). */
int tpt_hip_scene_create(const tpt_scene_desc *desc, tpt_hip_scene **out);
/* Frees the device arena (cuda::Scene::~Scene, cuda/scene.h:176-197). */
void tpt_hip_scene_destroy(tpt_hip_scene *scene);
/* Size in bytes of the device arena and of its BVH parts; `node_bytes` is the
 * size of one BVH node record (64 for the default quantized 4-wide layout).
 * Any output pointer may be NULL. */
int tpt_hip_scene_info(const tpt_hip_scene *scene, uint64_t *arena_bytes, uint64_t *num_nodes,
                       uint64_t *num_triangles, uint64_t *num_analytic, uint64_t *node_bytes);

/* PathTracer(hip).render(scene, spp): synchronous; writes H*W*3 floats (row 0
 * at the bottom) into the HOST buffer `out_rgb`.  Replaces
 * tinypt::PathTracer::render (path_tracer.h:26-44).  `stats` may be NULL. */
int tpt_hip_render(tpt_hip_scene *scene, const tpt_render_params *params, float *out_rgb,
                   tpt_render_stats *stats);

/* Render without allocation.  No counterpart in the reference; for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  Renders as tpt_hip_render does (the
 * same image for the same params and seed) with one path pool whose path state,
 * queues, counters and per-pixel sums live in a fixed device array of the
 * library (about 130 MB, 2^19 paths and 2^19 pixels; one array per device, held
 * until the process ends).  Enqueued on `stream` (a hipStream_t, NULL = the null
 * stream) with no synchronisation, no allocation and no memset: writes H*W*3
 * floats (row 0 at the bottom) into the DEVICE buffer `out_rgb`.  Rejects scenes
 * of more than 2^19 pixels, queue_capacity above 2^19 and count_traffic.  All
 * calls on one device share its array: calls that overlap on a device race
 * (order them on one stream). */
int tpt_hip_render_noalloc(tpt_hip_scene *scene, const tpt_render_params *params, float *out_rgb, void *stream);

/* ---- batch: many independent renders in one pipeline ----------------------
 * No counterpart in the reference, which renders one scene per
 * PathTracer::render call (path_tracer.h:26-44); each scene of a batch renders
 * exactly as tpt_hip_render would render it alone (same image, bit for bit). */
typedef struct {
    int32_t spp;   /* samples per pixel of this scene */
    int32_t pad_;
    uint64_t seed; /* stream seed of this scene */
} tpt_batch_item;
typedef struct {
    double render_ms;       /* wall time of the batch render, after tuning */
    double tune_ms;         /* wall time spent by the autotuner (0 when both sizes were given) */
    uint64_t paths;         /* total paths of the batch */
    int32_t queue_capacity; /* path slots per pipeline wave (tuned or given) */
    int32_t block_size;     /* threads per workgroup of generate / shade (tuned or given) */
    int32_t waves;          /* pipeline waves: batched waves + plain waves of large scenes */
    int32_t graphs;         /* 1: the render ran as one captured hipGraph launch; 0: per-wave graph replays */
} tpt_batch_stats;
/* Renders n scenes (all created on the current device with the same node layout, n
 * <= 1024) through one pipeline: pipeline waves (generate, then max_depth + 1 x
 * extend / shade / shadow) of up to queue_capacity paths, mixing scenes,
 * alternate between up to three path pools on as many streams (one with
 * TPT_STREAMS=1); each wave's accumulation follows the previous wave's, so the
 * images are bit-identical to single-scene renders.  One captured wave per pool
 * is replayed from the host (TPT_BATCH_GRAPH=1: the whole render is captured into
 * one hipGraph and launched once).
 * Scenes of at least 2^23 paths (TPT_BATCH_LARGE=10..40 sets the exponent) fill
 * the GPU alone: they run after the batched ones as plain pipeline waves of their
 * own (tpt_hip_render's kernels and 2^24-slot waves) on the same pools.
 * `params` gives the shared settings (max_depth, rr_depth, shadow_bounded, output,
 * accum_fp64); its spp and seed are ignored (per scene in `items`).
 * params->queue_capacity == 0 selects the capacity by autotuning over 2^18..2^22
 * slots (TPT_BATCH_MAX_CAP=23..26 raises the top), params->block_size == 0 the
 * workgroup size over 64..512.  The path pools stay allocated on the first
 * (fewest-paths) scene's handle for later batches.  out_rgb[i]
 * receives scene i's image (H*W*3 floats, row 0 at the bottom); `scene_ms` (may be
 * NULL) receives n completion times in ms since the batch started; `stats` may be
 * NULL. */
int tpt_hip_render_batch(tpt_hip_scene *const *scenes, const tpt_batch_item *items, int32_t n,
                         const tpt_render_params *params, float *const *out_rgb, tpt_batch_stats *stats,
                         double *scene_ms);

/* ---- learned radiance cache --------------------------------------------
 * No counterpart in the reference.  A 5-layer, 64-wide ReLU MLP (BF16 MFMA,
 * FP32 accumulate) that predicts the radiance a path vertex scatters towards
 * the camera, trained online (Adam) on path suffixes the tracer produces.  The
 * network belongs to the scene handle and keeps training across calls. */
typedef struct {
    int32_t frames;        /* progressive frames of one sample per pixel */
    int32_t cache_depth;   /* path-vertex depth at which paths terminate into the cache (2 = after two
                              bounces); < 0: no cache (plain progressive rendering) */
    int32_t train_records; /* training path suffixes per frame (1 .. 2^20) */
    int32_t reset;         /* 1: re-initialise the network from weight_seed before rendering */
    float learning_rate;   /* Adam step size */
    int32_t train_steps;   /* Adam steps per frame, each on 1 / train_steps of the frame's records (>= 1) */
    uint64_t weight_seed;
    int32_t suffix_len;    /* > 0: a training path also terminates into the cache, suffix_len vertices
                              after its cache vertex (self-training: its target then includes the cache's
                              prediction there); 0: training paths run to max_depth */
    int32_t pad_;
} tpt_cache_params;
typedef struct {
    double render_ms;      /* wall time of the render */
    double infer_ms;       /* device time of the MLP inference launches */
    double train_ms;       /* device time of the MLP training + Adam launches */
    double last_loss;      /* mean relative-L2 training loss of the last frame group's steps */
    uint64_t queries;      /* paths terminated into the cache per frame of the last frame group */
    uint64_t records;      /* training records per frame of the last frame group */
    int32_t steps;         /* Adam steps taken so far by this scene's network */
    int32_t pad_;
} tpt_cache_stats;
/* Progressive render: `cache->frames` frames of one sample per pixel (frame f
 * uses sample index f of the per-(pixel, sample) streams of params->seed); after
 * each frame the cache queries are evaluated, the frame's training records
 * become targets and the network takes train_steps Adam steps (environment
 * TPT_FRAME_GROUP = 2..4 traces that many frames per pipeline wave, sharing the
 * weights of the training before them; cache_depth < 0 groups frames up to about
 * 2^22 paths by default, which leaves the image unchanged).  Writes the HDR mean of
 * the frames (H*W*3 floats, row 0 at the bottom); params->spp and output are
 * ignored, accum_fp64 selects FP64 accumulation.  `stats` may be NULL. */
int tpt_hip_render_cached(tpt_hip_scene *scene, const tpt_render_params *params, const tpt_cache_params *cache,
                          float *out_rgb, tpt_cache_stats *stats);
/* Radiance-cache test hooks.  Weights: 5 x 64 x 64 FP32, [layer][in][out] (the
 * last layer uses out < 3); `set` / `get` may be NULL.  Setting weights resets
 * the Adam state. */
int tpt_hip_cache_weights(tpt_hip_scene *scene, const float *set, float *get);
/* Evaluates the network on n host records of 64 BF16 inputs (uint16 bits):
 * out = n x 3 floats. */
int tpt_hip_cache_eval(tpt_hip_scene *scene, const uint16_t *inputs, int32_t n, float *out);
/* One training step on n host records (inputs n x 64 BF16, targets n x 3
 * floats): gradient of the mean relative-L2 loss, then Adam with step size lr.
 * `loss` (may be NULL) receives the loss before the step; `grad` (may be NULL)
 * receives the gradient (5 x 64 x 64 FP32, [layer][in][out]). */
int tpt_hip_cache_train(tpt_hip_scene *scene, const uint16_t *inputs, const float *targets, int32_t n, float lr,
                        float *loss, float *grad);

/* ---- progressive sessions ------------------------------------------------
 * No counterpart in the reference, whose PathTracer::render (path_tracer.h:26-44)
 * is one synchronous call.  A session accumulates frames of one sample per pixel
 * on the device.  Frame f is sample f of the per-(pixel, sample) streams of
 * params->seed, so F frames added over any number of calls give the image one
 * tpt_hip_render_cached call of F frames gives (bit for bit without the cache;
 * with it, up to the order of the gradient sums).  `cache` may be NULL (plain
 * progressive rendering; its `frames` field is not used).  add only enqueues
 * the frames on the scene's stream and returns; wait and read block until they
 * are done; read writes the HDR mean of the frames so far (row 0 at the bottom).
 * The session owns its path pool; with a cache it trains the scene handle's
 * network.  Frames run on the scene handle's streams (frame f on lane
 * f % TPT_STREAMS); the training steps of every session and cached render of one
 * scene handle run one after another, and the cache hooks (weights, eval, train)
 * wait for all of the handle's streams first. */
typedef struct tpt_hip_progress tpt_hip_progress;
int tpt_hip_progress_create(tpt_hip_scene *scene, const tpt_render_params *params, const tpt_cache_params *cache,
                            tpt_hip_progress **out);
int tpt_hip_progress_add(tpt_hip_progress *session, int32_t frames);
int tpt_hip_progress_wait(tpt_hip_progress *session);
int tpt_hip_progress_read(tpt_hip_progress *session, float *out_rgb, int32_t *frames);
void tpt_hip_progress_destroy(tpt_hip_progress *session);

/* Closest hit for n rays of 8 floats (org.xyz, tmin, dir.xyz, tmax); directions
 * must be unit length as in the reference.  Writes the global primitive id
 * (-1 on miss) and hit distance (INF on miss).  Geometric test only (no alpha
 * cut-out).  `flags & TPT_PTR_DEVICE`: all pointers are device memory and the
 * call is enqueued on `stream` (a hipStream_t, NULL = default) without a sync. */
#define TPT_PTR_DEVICE 1u
#define TPT_COUNT_TRAFFIC 2u
int tpt_hip_trace_closest(tpt_hip_scene *scene, const float *rays, uint64_t n, int32_t *prim_id, float *t,
                          uint32_t flags, void *stream);
/* Any hit with tmin <= t < tmax: writes 1 when occluded, else 0. */
int tpt_hip_trace_occluded(tpt_hip_scene *scene, const float *rays, uint64_t n, uint8_t *occluded, uint32_t flags,
                           void *stream);
/* Node / triangle fetch counters of the last TPT_COUNT_TRAFFIC trace call. */
int tpt_hip_last_trace_counters(tpt_hip_scene *scene, uint64_t *nodes_visited, uint64_t *tris_tested);

#ifdef __cplusplus
}
#endif
#endif /* TINYPT_HIP_H_ */
