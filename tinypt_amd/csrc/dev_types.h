// dev_types.h — POD layouts of the device arena, shared by the host scene
// builder (scene_build.cpp) and the gfx950 kernels (kernels.hip).
//
// Arena layout (one hipMalloc per scene, every section 256-B aligned):
//   nodes      one of three node layouts (DScene.layout):
//              kNodeBvh4Q  4-wide, 64 B (default): origin.xyz, scale.xyz (powers of two), 8-bit child
//                          bounds qlo_x[4] qhi_x[4] qlo_y[4] qhi_y[4] qlo_z[4] qhi_z[4], child refs[4];
//                          a child box is origin + q * scale, rounded outwards at build time
//              kNodeBvh4   4-wide, 128 B: f32 child boxes lo_x[4] hi_x[4] lo_y[4] hi_y[4] lo_z[4] hi_z[4],
//                          child refs[4], pad
//              kNodeBvh2   binary, 64 B: both child boxes (f32) + refs
//   tris_bvh   triangles in BVH leaf order, 48 B each: (a.xyz, gid) (ab.xyz, tri_id) (ac.xyz, flags)
//   tris       the same records in triangle-id order (surface / light-pdf lookups)
//   tri_info   2 x int4 per triangle: (material, na, nb, nc) (ta, tb, tc, mesh)
//   shade_rec  128 B per triangle, one cache line: everything shading reads about a hit
//              triangle (edges, material, vertex normals, texcoords), see kShadeRec
//   normals, texcoords, gid_local, analytic, materials, images, texels, lights, env CDF
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

namespace tpt {

// bit casts for host code that fills device records
inline float __int_as_float_host(int32_t i) {
    float f;
    memcpy(&f, &i, 4);
    return f;
}
inline int32_t __float_as_int_host(float f) {
    int32_t i;
    memcpy(&i, &f, 4);
    return i;
}

static constexpr float kINF = 1e20f; // cpu/defs.h:23
static constexpr float kEPS = 1e-3f; // cpu/defs.h:24

// child reference inside a node: >= 0 inner node index; < 0 leaf
//   leaf = 0x80000000 | (first_tri << 3) | (count - 1), count <= 8
// The empty slots of a 4-wide node are leaves of one zero-size sentinel triangle
// (the last leaf-order record), which no ray hits.
static constexpr uint32_t kLeafBit = 0x80000000u;
static constexpr int kMaxLeaf = 8;
static constexpr int kStackSize = 96; // worst-case traversal stack entries (checked at BVH build)
static constexpr int kLdsStack = 16;  // entries per lane kept in LDS
static constexpr int kLdsRow = kLdsStack + 1; // LDS words per lane (odd: conflict-free rows)
enum { kNodeBvh2 = 0, kNodeBvh4 = 1, kNodeBvh4Q = 2 };
static constexpr int kNodeFloat4[3] = {4, 8, 4}; // float4 records per node, by layout

struct DTex {
    int32_t image;   // -1: constant
    int32_t channel; // first channel
    float v[3];
};

struct DImage {
    int32_t w, h, c, fmt; // fmt 0 = f32, 1 = u8
    uint64_t offset;      // byte offset into the texel blob
};

struct DLobe {
    int32_t type;
    int32_t specular;
    float param;  // ior / shininess
    float alpha2; // GGX alpha^2 = 2/(s+2)
    DTex color;
};

struct DMaterial {
    DLobe lobe[2];
    int32_t combine;
    float weight;
    int32_t specular;     // surface()->is_specular()
    int32_t alpha_opaque; // AlphaTexture::empty()
    DTex emission;
    DTex alpha;
    int32_t bump;         // gradient image or -1
    int32_t emissive;
};

enum { A_SPHERE = 0, A_RECT = 1, A_CIRCLE = 2 };
struct DAnalytic {
    int32_t type, material, gid, pad;
    float center[3];
    float radius;
    float u[3], v[3], w[3]; // frame columns
    float dim[2];
    float dist;
    float pad2;
};

struct DDelta {
    int32_t type; // 0 distant, 1 point
    float u[3], v[3], w[3];
    float center[3];
    float emission[3]; // color * power
    float radius;      // point radius
    float angular_radius;
};

// L_MESH: ObjectLight of a whole mesh: `local` = first of its `gid` shapes in
// DScene::light_shapes (tri_id ranges), sampled shape-uniform then face-uniform
enum { L_ANALYTIC = 0, L_TRIANGLE = 1, L_ENV = 2, L_MESH = 3 };
struct DLight {
    int32_t kind;
    int32_t local; // analytic index / triangle id
    int32_t gid;
    int32_t pad;
};

struct DCamera {
    float center[3];
    float u[3], v[3], w[3];
    float direct[3], right[3], up[3];
    int32_t width, height;
    float aperture;
    float pad;
};

// Shading record of a triangle, kShadeRec float4 (128 B, one L2 line):
//   [0] ab.xyz, material   [1] ac.xyz, flags (1: vertex normals, 2: texcoords)
//   [2] na.xyz, nb.x       [3] nb.yz, nc.xy   [4] nc.z, ta.xy, tb.x   [5] tb.y, tc.xy, 0
//   [6], [7] unused
static constexpr int kShadeRec = 8;

struct DScene {
    const float4 *nodes;
    const float4 *tris_bvh;
    const float4 *tris;
    const int4 *tri_info;
    const float4 *shade_rec; // kShadeRec per triangle, in triangle-id order
    const float *normals;
    const float *texcoords;
    const int32_t *gid_local; // analytic: -(idx + 1); triangle: tri_id
    const DAnalytic *analytic;
    const DMaterial *materials;
    const DImage *images;
    const uint8_t *texels;
    const DDelta *delta;
    const DLight *lights;
    const float *env_marg; // env_h + 1 floats
    const float *env_cond; // env_h * (env_w + 1) floats
    // guide tables of the CDFs (env_gh + 1 and env_h * (env_gw + 1) entries): entry k of
    // a table is upper_bound(cdf, k / G) for k < G, and count + 1 for k = G
    const int32_t *env_gmarg;
    const int32_t *env_gcond;
    const int2 *light_shapes; // [first, end) tri_id range of each shape of the L_MESH lights
    int32_t num_nodes;
    int32_t root; // child reference of the root (inner 0, or a leaf when the BVH is one leaf), -1 if empty
    int32_t num_tris;
    int32_t num_analytic;
    int32_t num_delta;
    int32_t num_lights;
    int32_t env_w, env_h;
    int32_t bg_map;
    int32_t layout;     // kNodeBvh2 / kNodeBvh4 / kNodeBvh4Q
    int32_t persistent; // 0: one ray per lane until the whole wave is done; 1: persistent kernels with
                        // per-wave ray refill; 2: persistent + parked leaves; 3: each wave traces its own
                        // `chunk` consecutive queue entries, refilling idle lanes from them; 4: 3 + parked leaves
    int32_t num_cu;     // compute units of the device (host-side grid sizing)
    int32_t lds_stack;  // 1: first kLdsStack traversal-stack entries in LDS; 0: private (scratch) stack
    int32_t lds_nodes;  // quantized layout: nodes [0, lds_nodes) (the top of the tree in BFS order)
                        // are copied into LDS by each traversal workgroup; 0 = no node cache
    int32_t trav_block; // threads per workgroup of the persistent traversal kernels
    int32_t max_lds;    // bytes of LDS one workgroup may allocate on the device
    int32_t chunk;      // queue entries per wave of the chunked traversal kernels (persistent 3, 4)
    int32_t env_gh;     // guide buckets of the marginal CDF (a power of two)
    int32_t env_gw;     // guide buckets of each conditional CDF (a power of two)
    int32_t loc_bits;   // > 0: a triangle's location (leaf records, hit records) is
                        // tri_id | material << loc_bits, so shading knows the material
                        // with the hit; 0: the location is the tri_id alone
    DTex bg;
    float bg_rot[9]; // columns u, v, w
    DCamera cam;
};

// Per-render parameters handed to the kernels.
struct DRender {
    uint64_t seed;
    int32_t spp_eff;  // n_sub * b * b
    int32_t block;    // b
    int32_t nsub;
    int32_t max_depth;
    int32_t rr_depth;
    int32_t shadow_bounded;
    int32_t output;
    int32_t accum_fp64;
    int32_t pixels;   // W * H
    int32_t k;        // samples per pixel in one wave
};

// Wavefront path state, SoA over slots (slot = s_local * pixels + pixel).
struct DPaths {
    float4 *org;   // xyz origin, w unused
    float4 *dir;   // xyz direction
    float4 *decay; // xyz throughput
    float4 *color; // xyz radiance
    uint64_t *rng; // 48-bit erand48 state
    uint64_t *key; // stream seed (alpha hash key)
    int32_t *depth;
    // queues
    int32_t *queue_a;
    int32_t *queue_b;
    int32_t *shadow_slot;
    float4 *shadow_org;  // xyz, w = tmax
    float4 *shadow_dir;  // xyz
    float4 *shadow_val;  // contribution
    float4 *hit;         // t, gid(bits), u, v — indexed by queue position
    uint32_t *counters;  // kCntWords words, see kCntQueue
    // per pixel accumulation
    float *sub_f;   // 3 per pixel: running sub-pixel sum (fp32)
    double *sub_d;  // 3 per pixel (fp64 mode)
    float *pix_f;   // 3 per pixel: sum of clipped sub-pixel means (or hdr sum)
    double *pix_d;  // 3 per pixel (fp64 hdr)
    int32_t *sid;   // batch mode: scene of each slot
};

// Words of DPaths.counters.  Every counter that kernels bump with atomics sits
// on its own 256-B line: lanes read the queue counts while other waves append
// to the next queue, and a shared line would queue those reads behind the
// atomics at one L2 channel.
enum : int {
    kCntLine = 64,             // u32 words per 256-B line
    kCntQueue = 0,             // ray-queue counts: queue q (0 queue_a, 1 queue_b) at q * kCntLine
    kCntShadow = 2 * kCntLine, // shadow-queue count
    kCntFetch = 3 * kCntLine,  // refill counter of the persistent extend kernel
    kCntShFetch = 4 * kCntLine, // refill counter of the persistent shadow kernel
    kCntTraffic = 5 * kCntLine, // 4 x u64: nodes visited, triangles tested, closest rays, shadow rays
    kCntWords = 6 * kCntLine
};

// Learned radiance cache hooks of the shade kernel.  At the path vertex of depth
// `depth`, a path either terminates into the cache (a query: its encoded vertex
// is evaluated by the MLP afterwards and color += decay * prediction), or, for
// one path in `train_stride`, is recorded as a training example (its throughput
// T and radiance C at the vertex) and continues; its target is then
// (final radiance - C) / T, the radiance the vertex scattered towards the camera.
struct DCache {
    int32_t depth;        // cache vertex depth (< 0: cache off)
    int32_t train_stride; // one path in train_stride (by a per-frame hash of its slot) trains
    uint32_t frame;       // frame index (training-path hash)
    int32_t train_cap;    // capacity of the training-record list
    float lo[3];          // scene bounds: position encoding (p - lo) * inv_ext
    float inv_ext[3];
    int32_t suffix;       // > 0: training paths terminate into the cache `suffix` vertices after their cache vertex
    float min_rough;      // a path ends in the cache at its first vertex of depth >= `depth` whose first
                          // lobe's roughness is >= min_rough (0: at depth `depth`, whatever the material)
    int32_t albedo_factor; // the network predicts the vertex's scattered radiance divided by its first
                           // lobe's albedo (queries multiply it back), so textures need not be learned:
                           // 1 where that lobe's colour is a texture map, 2 at every vertex, 0 nowhere
    float light_prob;      // probability of sampling the light mixture at a vertex before the cache
    uint16_t *feat;       // per slot: 64 BF16 inputs of the vertex where the path queries the cache
    uint16_t *feat_t;     // per slot: 64 BF16 inputs of a training path's cache vertex
    float4 *rec_T;        // per slot: throughput arriving at the vertex, w = its depth (-1: none yet)
                          // (training paths)
    float4 *rec_C;        // per slot: radiance after the vertex's emission (training paths)
    int32_t *q_slot;      // query list (slots), count in counters[0]
    int32_t *t_slot;      // training list (slots), count in counters[1]
    uint32_t *counters;
};

// Batch mode: many independent scenes share one path pool and one pipeline.
// Global path index g runs scene-major (scene 0's paths first), and within a
// scene sample-major (g - path_off[s] = sample * pixels + pixel), so every
// pixel's samples arrive in sample order and the accumulation order equals a
// single-scene render's.  A pipeline wave covers global paths [g0, g0 + cap).
// Ray and shadow queues are split into one contiguous segment per scene, so a
// wavefront of the persistent traversal kernels works on one scene at a time.
enum {
    kSegQaBase = 0, kSegQaCnt, kSegQbBase, kSegQbCnt, // ping-pong ray queues
    kSegFetch,                                        // extend refill counter
    kSegShBase, kSegShCnt, kSegShFetch,               // shadow queue
    kSegRayChunk, kSegShChunk,                        // chunked traversal: exclusive prefix of chunks per
                                                      // scene (ray / shadow queue), total at index n, the
                                                      // chunk size at n + 1
    kSegFields
};
struct DBatch {
    const DScene *scenes;       // n; array pointers stored as byte offsets from `base`
    const uint8_t *base;        // lowest scene arena of the batch
    const DRender *renders;     // n: per-scene sampler (spp_eff, block, nsub, pixels) + shared depth / RR / output
    const uint64_t *path_off;   // n + 1
    const uint64_t *pix_off;    // n + 1: first pixel of each scene in the concatenated accumulators / output
    const uint64_t *seed_mixed; // n: splitmix64(seed) of each scene
    int32_t *seg;               // kSegFields * seg_stride, field-major: seg[f * seg_stride + s]
    uint64_t *state;            // [0] g0 of the current pipeline wave, [1] total paths
    int32_t n;
    int32_t cap;
    int32_t seg_stride;         // n + 2 rounded up to kSegAlign: every field starts on its own cache lines
    int32_t pools;              // pipeline waves in flight (path pools); a pool's next wave starts pools * cap later
};
// Segment fields are kept on separate 256-B lines: the counters that a kernel
// bumps with atomics would otherwise share a line with the bases and counts
// that every lane of the same kernel reads, and those reads would queue behind
// the atomics at one L2 channel.
static constexpr int kSegAlign = 64;
inline int seg_stride_for(int n) { return (n + 2 + kSegAlign - 1) / kSegAlign * kSegAlign; }

} // namespace tpt
