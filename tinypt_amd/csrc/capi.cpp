// capi.cpp — the extern "C" boundary of libtinypt_hip.so (declared in
// include/tinypt_hip.h) and the host side of the wavefront render loop.
//
// Replaces, for device=hip, tinypt::Scene::to(Device) (li-plus/tinypt
// src/tinypt/scene.h:24-37), cuda::Scene (cuda/scene.h:95-197),
// tinypt::PathTracer::render (path_tracer.h:26-44) and the CUDA launch
// (cuda/path_tracer_kernel.cu:115-120).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <chrono>
#include <cstring>
#include <mutex>
#include <numeric>
#include <type_traits>
#include <string>
#include <vector>

#include "../../include/tinypt_hip.h"
#include "dev_math.h"
#include "dev_types.h"
#include "nrc.h"
#include "scene_build.h"

namespace tpt {
void launch_generate(const DScene &, const DRender &, const DPaths &, int, int, uint64_t, int, hipStream_t);
void launch_megapath(const DScene &, const DRender &, const DPaths &, int, int, uint64_t, int, int, int, hipStream_t);
void launch_extend(const DScene &, const DPaths &, const int32_t *, int, int, int, int, bool, uint64_t, hipStream_t);
void launch_shade(const DScene &, const DRender &, const DPaths &, const int32_t *, int32_t *, int, int, int, int,
                  hipStream_t, const DCache *);
void launch_shadow(const DScene &, const DPaths &, int, int, bool, uint64_t, hipStream_t);
void launch_accumulate(const DRender &, const DPaths &, int, int, int, hipStream_t);
void launch_finalize(const DRender &, const DPaths &, float *, hipStream_t);
void launch_trace_closest(const DScene &, const float4 *, uint64_t, int32_t *, float *, uint32_t *, bool, int, int,
                          hipStream_t);
void launch_trace_occluded(const DScene &, const float4 *, uint64_t, uint8_t *, uint32_t *, int, int, hipStream_t);
void launch_batch_wave(const DBatch &, const DPaths &, int, int, int, int, bool, int, int, int, int, uint64_t,
                       hipStream_t);
void launch_batch_tail(const DBatch &, const DPaths &, uint64_t, hipStream_t);
void launch_batch_zero(const DBatch &, const DPaths &, int, uint64_t, hipStream_t);
void launch_batch_state(const DBatch &, uint64_t, uint64_t, hipStream_t);
void launch_stamp(uint64_t *, hipStream_t);
void launch_batch_finalize(const DBatch &, const DPaths &, float *, uint64_t, hipStream_t);
void launch_cache_reset(const DCache &, hipStream_t);
void launch_cache_apply(const DPaths &, const DCache &, const float4 *, int, hipStream_t);
void launch_cache_targets(const DPaths &, const DCache &, float4 *, hipStream_t);
void launch_frame_accumulate(const DPaths &, int, int, int, int, hipStream_t);
void launch_frame_finalize(const DPaths &, int, int, int, float *, hipStream_t);
void launch_fill_u32(uint32_t *, int, uint32_t, hipStream_t);
} // namespace tpt

using namespace tpt;

namespace {

thread_local std::string g_err;

int set_err(int code, const std::string &m) {
    g_err = m;
    return code;
}

#define HIP_OK(expr)                                                                                                   \
    do {                                                                                                               \
        hipError_t e_ = (expr);                                                                                        \
        if (e_ != hipSuccess)                                                                                          \
            throw tpt_error(TPT_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_) + " (" __FILE__ ":" +      \
                                             std::to_string(__LINE__) + ")");                                          \
    } while (0)

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// Integer tuning knob from the environment, clamped to [lo, hi]; `def` when unset.
int env_int(const char *name, int def, int lo, int hi) {
    const char *v = std::getenv(name);
    if (!v || !*v) return def;
    return std::min(hi, std::max(lo, std::atoi(v)));
}

// Defaults of the persistent traversal kernels: workgroup size and the number of
// top-of-tree nodes (64 B each) each workgroup stages in LDS.
constexpr int kTravBlock = 512;
constexpr int kLdsNodes = 512;

struct PathBuffers {
    void *mem = nullptr;
    size_t bytes = 0;
    size_t slots = 0;
    int pixels = 0;
    DPaths P{};
    float *out = nullptr;
};

// Per-slot staging buffers of one cached frame in flight (one set per path pool).
struct CacheBufs {
    void *bmem = nullptr; // per-slot buffers for `slots` paths
    size_t slots = 0;
    int train_cap = 0;
    uint16_t *feat = nullptr, *feat_t = nullptr;
    float4 *rec_T = nullptr, *rec_C = nullptr, *pred = nullptr, *target = nullptr;
    int32_t *q_slot = nullptr, *t_slot = nullptr;
    uint32_t *counters = nullptr; // [0] queries, [1] training records
    ~CacheBufs() {
        if (bmem) (void)hipFree(bmem);
    }
};

// Radiance-cache state of a scene handle: the network (weights and Adam state,
// kept across renders: the cache trains online) and the staging buffers.
struct CacheState {
    void *wmem = nullptr; // weights, moments, gradient, BF16 copies, loss
    NrcWeights W{};
    int step = 0;
    CacheBufs buf[4]; // one set per frame lane
    ~CacheState() {
        if (wmem) (void)hipFree(wmem);
    }
};

} // namespace

struct tpt_hip_scene {
    int device = 0;
    int cus = 256; // compute units of the device (grid sizing)
    void *arena = nullptr;
    size_t arena_bytes = 0;
    DScene S{};
    uint64_t num_nodes = 0, num_tris = 0, num_analytic = 0;
    PathBuffers pb;
    PathBuffers pb2;                    // path pool of lane 1 (several-stream renders), pb's accumulators
    uint32_t *trace_counters = nullptr; // 8 x u32
    uint64_t last_nodes = 0, last_tris = 0;
    hipStream_t stream = nullptr;
    hipStream_t stream2 = nullptr;      // stream of lane 1 (several-stream renders), default priority
    hipStream_t streamx[2] = {nullptr, nullptr}; // streams of lanes 2, 3
    PathBuffers pbx[2];                 // path pools of lanes 2, 3, pb's accumulators
    PathBuffers batch_pb[4];            // path pools of batches whose first (shortest) scene this is,
                                        // kept across calls (with the per-slot scene ids)
    int streams = 3;                    // plain renders alternate pipeline waves, cached and progressive
                                        // renders frames, over `streams` path pools on as many streams;
                                        // batches use up to three pools
    float bounds_lo[3] = {0, 0, 0}, bounds_hi[3] = {1, 1, 1}; // finite geometry (radiance-cache encoding)
    bool env_light = false; // the light mixture holds the environment (cached frames' light sampling)
    CacheState cache;
    // recorded after the latest training step of the scene's network, whichever
    // cached render or session enqueued it: every inference and training waits on
    // it, so the network's steps never overlap (they share the gradient buffer)
    hipEvent_t train_ev = nullptr;
    bool train_any = false;
    ~tpt_hip_scene() {
        if (stream) (void)hipStreamSynchronize(stream);
        if (stream2) (void)hipStreamSynchronize(stream2);
        for (hipStream_t x : streamx)
            if (x) (void)hipStreamSynchronize(x);
        if (pb.mem) (void)hipFree(pb.mem);
        if (pb2.mem) (void)hipFree(pb2.mem);
        for (PathBuffers &x : pbx)
            if (x.mem) (void)hipFree(x.mem);
        for (PathBuffers &x : batch_pb)
            if (x.mem) (void)hipFree(x.mem);
        if (stream2) (void)hipStreamDestroy(stream2);
        for (hipStream_t x : streamx)
            if (x) (void)hipStreamDestroy(x);
        if (arena) (void)hipFree(arena);
        if (trace_counters) (void)hipFree(trace_counters);
        if (train_ev) (void)hipEventDestroy(train_ev);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

template <typename T>
size_t put(std::vector<uint8_t> &blob, const std::vector<T> &v) {
    size_t off = align256(blob.size());
    blob.resize(off + v.size() * sizeof(T));
    if (!v.empty()) std::memcpy(blob.data() + off, v.data(), v.size() * sizeof(T));
    return off;
}

void upload_scene(tpt_hip_scene *h, const HostScene &H) {
    std::vector<uint8_t> blob;
    size_t o_nodes = put(blob, H.nodes);
    size_t o_tb = put(blob, H.tris_bvh);
    size_t o_t = put(blob, H.tris);
    size_t o_ti = put(blob, H.tri_info);
    size_t o_sr = put(blob, H.shade_rec);
    size_t o_n = put(blob, H.normals);
    size_t o_tc = put(blob, H.texcoords);
    size_t o_gl = put(blob, H.gid_local);
    std::vector<DAnalytic> an = H.analytic;
    if (an.empty()) an.resize(1);
    size_t o_an = put(blob, an);
    std::vector<DMaterial> mats = H.materials;
    if (mats.empty()) mats.resize(1);
    size_t o_m = put(blob, mats);
    std::vector<DImage> imgs = H.images;
    if (imgs.empty()) imgs.resize(1);
    size_t o_im = put(blob, imgs);
    size_t o_tx = put(blob, H.texels);
    std::vector<DDelta> dl = H.delta;
    if (dl.empty()) dl.resize(1);
    size_t o_dl = put(blob, dl);
    std::vector<DLight> li = H.lights;
    if (li.empty()) li.resize(1);
    size_t o_li = put(blob, li);
    size_t o_em = put(blob, H.env_marg);
    size_t o_ec = put(blob, H.env_cond);
    size_t o_gm = put(blob, H.env_gmarg);
    size_t o_gc = put(blob, H.env_gcond);
    std::vector<int2> ls = H.light_shapes;
    if (ls.empty()) ls.resize(1);
    size_t o_ls = put(blob, ls);
    size_t total = align256(blob.size());
    blob.resize(total);
    HIP_OK(hipMalloc(&h->arena, total));
    h->arena_bytes = total;
    HIP_OK(hipMemcpy(h->arena, blob.data(), total, hipMemcpyHostToDevice));
    uint8_t *base = (uint8_t *)h->arena;
    DScene &S = h->S;
    S.nodes = (const float4 *)(base + o_nodes);
    S.tris_bvh = (const float4 *)(base + o_tb);
    S.tris = (const float4 *)(base + o_t);
    S.tri_info = (const int4 *)(base + o_ti);
    S.shade_rec = (const float4 *)(base + o_sr);
    S.normals = (const float *)(base + o_n);
    S.texcoords = (const float *)(base + o_tc);
    S.gid_local = (const int32_t *)(base + o_gl);
    S.analytic = (const DAnalytic *)(base + o_an);
    S.materials = (const DMaterial *)(base + o_m);
    S.images = (const DImage *)(base + o_im);
    S.texels = (const uint8_t *)(base + o_tx);
    S.delta = (const DDelta *)(base + o_dl);
    S.lights = (const DLight *)(base + o_li);
    S.env_marg = (const float *)(base + o_em);
    S.env_cond = (const float *)(base + o_ec);
    S.env_gmarg = (const int32_t *)(base + o_gm);
    S.env_gcond = (const int32_t *)(base + o_gc);
    S.light_shapes = (const int2 *)(base + o_ls);
    S.layout = H.layout;
    S.loc_bits = H.loc_bits;
    const char *pers = std::getenv("TPT_PERSISTENT");
    // traversal kernels (room 1024x768x256, Mpaths/s):
    //   0 one ray per lane; a wave keeps its 64 consecutive queue entries until all are done: 331
    //   1 persistent, idle lanes refilled from one launch-wide counter: 237
    //   2 1 + parked leaves: 237
    //   3 each wave refills from its own chunk of consecutive queue entries (no shared counter): 414
    //   4 (default) 3 + parked leaves: 434
    // Neighbouring queue entries are mostly neighbouring pixels, so a wave whose rays stay
    // neighbours fetches the same nodes; the launch-wide counter both mixes rays from all over
    // the queue and serialises every refill at one L2 channel (L2 read latency 700 vs 233 cycles).
    S.persistent = 4;
    if (pers && std::string(pers).size() == 1 && pers[0] >= '0' && pers[0] <= '4') S.persistent = pers[0] - '0';
    // queue entries per wavefront of the chunked traversal.  Room 1024x768x256, depth 8:
    // chunks of 256 / 384 / 512 / 640 / 768 / 1024 / 1536 give three-stream frames of
    // 620-622 / 628-630 / 625-630 / 624-629 / 619-625 / 615 / 595 Mpaths/s and extend alone
    // 229 / 227 / 228 / 233 / 239 / 247 / 257 ms per frame; city 1920x1080x64: 512 and 768
    // give the same frame, extend alone 186 against 195 ms
    S.chunk = env_int("TPT_CHUNK", 512, 64, 1 << 20) / 64 * 64;
    // traversal stack: the first kLdsStack entries in LDS by default (the private
    // scratch stack spills to memory: ~9x the extend kernel's HBM writes);
    // TPT_LDS_STACK=0 selects the private stack
    const char *lds = std::getenv("TPT_LDS_STACK");
    S.lds_stack = (lds && std::string(lds) == "0") ? 0 : 1;
    // persistent traversal workgroups and the LDS node cache of the quantized layout
    // (the top of the tree, in BFS order at the start of the node array)
    S.trav_block = env_int("TPT_TRAV_BLOCK", kTravBlock, 64, 1024) / 64 * 64;
    // node cache: 512 nodes for the persistent kernels (one large workgroup per SIMD pair);
    // none for one-ray-per-lane (its 2048 resident threads per CU need the LDS for their stacks)
    const int nodes_default = (S.persistent == 1 || S.persistent == 2) ? kLdsNodes : 0;
    S.lds_nodes = H.layout == kNodeBvh4Q ? std::min(H.bfs_prefix, env_int("TPT_LDS_NODES", nodes_default, 0, 2048)) : 0;
    S.num_nodes = (int32_t)(H.nodes.size() / kNodeFloat4[H.layout]);
    S.root = H.root;
    S.num_tris = H.root == -1 ? 0 : (int32_t)(H.tris_bvh.size() / 3 - 1); // leaf-order records less the sentinel
    S.num_analytic = (int32_t)H.analytic.size();
    S.num_delta = (int32_t)H.delta.size();
    S.num_lights = (int32_t)H.lights.size();
    S.env_w = H.env_w;
    S.env_h = H.env_h;
    S.env_gw = H.env_gw;
    S.env_gh = H.env_gh;
    S.bg_map = H.bg_map;
    S.bg = H.bg;
    std::memcpy(S.bg_rot, H.bg_rot, sizeof(S.bg_rot));
    S.cam = H.cam;
    h->num_nodes = H.root >= 0 && !(H.root & (int32_t)kLeafBit) ? (uint64_t)S.num_nodes : 0;
    h->num_tris = H.root == -1 ? 0 : (uint64_t)S.num_tris;
    h->num_analytic = H.analytic.size();
    // bounds of the finite geometry: triangles, and analytic objects of radius /
    // extent below 1e4 (the radius-1e5 wall spheres of smallpt-style scenes are
    // left out so that the position encoding keeps its resolution)
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    auto grow = [&](float x, float y, float z, float r) {
        const float p[3] = {x, y, z};
        for (int a = 0; a < 3; a++) {
            lo[a] = std::min(lo[a], p[a] - r);
            hi[a] = std::max(hi[a], p[a] + r);
        }
    };
    if (H.root != -1)
        for (size_t t = 0; t + 2 < H.tris.size(); t += 3) {
            float4 a = H.tris[t], ab = H.tris[t + 1], ac = H.tris[t + 2];
            grow(a.x, a.y, a.z, 0.f);
            grow(a.x + ab.x, a.y + ab.y, a.z + ab.z, 0.f);
            grow(a.x + ac.x, a.y + ac.y, a.z + ac.z, 0.f);
        }
    for (const DAnalytic &an2 : H.analytic) {
        float r = an2.type == A_SPHERE || an2.type == A_CIRCLE ? an2.radius : std::max(an2.dim[0], an2.dim[1]);
        if (r < 1e4f) grow(an2.center[0], an2.center[1], an2.center[2], r);
    }
    for (int a = 0; a < 3; a++) {
        if (!(lo[a] <= hi[a])) {
            lo[a] = -1.f;
            hi[a] = 1.f;
        }
        h->bounds_lo[a] = lo[a];
        h->bounds_hi[a] = std::max(hi[a], lo[a] + 1e-3f);
    }
    h->env_light = false;
    for (const DLight &l : H.lights) h->env_light = h->env_light || l.kind == L_ENV;
}

// Byte offsets of a path pool of `slots` paths + accumulators / output of `pixels`
// pixels in one block (every array 256-B aligned).  `with_sid`: also the per-slot
// scene ids of batch mode; `with_out`: also the finished image.
struct PathLayout {
    size_t org, dir, dec, col, rng, key, dep, qa, qb, ss, so, sd, sv, hit, cnt, sf, sdd, pf, pd, out, sid, total;
};
constexpr size_t align256c(size_t x) { return (x + 255) & ~(size_t)255; }
constexpr PathLayout path_layout(size_t slots, size_t pixels, bool with_sid, bool with_out) {
    PathLayout L{};
    size_t off = 0;
    auto take = [&off](size_t bytes) {
        size_t o = off;
        off = align256c(off + bytes);
        return o;
    };
    L.org = take(slots * 16), L.dir = take(slots * 16), L.dec = take(slots * 16), L.col = take(slots * 16);
    L.rng = take(slots * 8), L.key = take(slots * 8), L.dep = take(slots * 4);
    L.qa = take(slots * 4), L.qb = take(slots * 4), L.ss = take(slots * 4);
    L.so = take(slots * 16), L.sd = take(slots * 16), L.sv = take(slots * 16), L.hit = take(slots * 16);
    L.cnt = take(kCntWords * 4);
    L.sf = take(pixels * 12), L.sdd = take(pixels * 24), L.pf = take(pixels * 12), L.pd = take(pixels * 24);
    L.out = with_out ? take(pixels * 12) : 0;
    L.sid = with_sid ? take(slots * 4) : 0;
    L.total = off;
    return L;
}

// Points P (and *out) into a block laid out by path_layout.
void bind_paths(uint8_t *b, const PathLayout &L, bool with_sid, bool with_out, DPaths &P, float **out) {
    P.sid = with_sid ? (int32_t *)(b + L.sid) : nullptr;
    P.org = (float4 *)(b + L.org);
    P.dir = (float4 *)(b + L.dir);
    P.decay = (float4 *)(b + L.dec);
    P.color = (float4 *)(b + L.col);
    P.rng = (uint64_t *)(b + L.rng);
    P.key = (uint64_t *)(b + L.key);
    P.depth = (int32_t *)(b + L.dep);
    P.queue_a = (int32_t *)(b + L.qa);
    P.queue_b = (int32_t *)(b + L.qb);
    P.shadow_slot = (int32_t *)(b + L.ss);
    P.shadow_org = (float4 *)(b + L.so);
    P.shadow_dir = (float4 *)(b + L.sd);
    P.shadow_val = (float4 *)(b + L.sv);
    P.hit = (float4 *)(b + L.hit);
    P.counters = (uint32_t *)(b + L.cnt);
    P.sub_f = (float *)(b + L.sf);
    P.sub_d = (double *)(b + L.sdd);
    P.pix_f = (float *)(b + L.pf);
    P.pix_d = (double *)(b + L.pd);
    if (out) *out = with_out ? (float *)(b + L.out) : nullptr;
}

// Path pool of `slots` paths + accumulators / output of `pixels` pixels (one
// allocation).  `with_sid`: also the per-slot scene ids of batch mode.
void alloc_paths(PathBuffers &pb, size_t slots, size_t pixels, bool with_sid) {
    if (pb.mem) {
        HIP_OK(hipFree(pb.mem));
        pb.mem = nullptr;
    }
    const PathLayout L = path_layout(slots, pixels, with_sid, true);
    HIP_OK(hipMalloc(&pb.mem, L.total));
    pb.bytes = L.total;
    pb.slots = slots;
    pb.pixels = (int)std::min<size_t>(pixels, (size_t)INT32_MAX);
    bind_paths((uint8_t *)pb.mem, L, with_sid, true, pb.P, &pb.out);
    HIP_OK(hipMemset(pb.P.counters, 0, kCntWords * 4));
}

void ensure_paths(tpt_hip_scene *h, size_t slots, int pixels) {
    PathBuffers &pb = h->pb;
    if (pb.mem && pb.slots >= slots && pb.pixels == pixels) return;
    if (pb.mem) {
        HIP_OK(hipStreamSynchronize(h->stream));
        HIP_OK(hipFree(pb.mem));
        pb.mem = nullptr;
    }
    alloc_paths(pb, slots, (size_t)pixels, false);
}

void ensure_stream2(tpt_hip_scene *h) {
    if (!h->stream2) HIP_OK(hipStreamCreateWithFlags(&h->stream2, hipStreamNonBlocking));
}

// Device memory owned by a call (freed on every exit).
struct DeviceBuffer {
    void *ptr = nullptr;
    explicit DeviceBuffer(size_t bytes) { HIP_OK(hipMalloc(&ptr, bytes)); }
    DeviceBuffer(const DeviceBuffer &) = delete;
    DeviceBuffer &operator=(const DeviceBuffer &) = delete;
    ~DeviceBuffer() {
        if (ptr) (void)hipFree(ptr);
    }
};
// An executable graph owned by a call.
struct GraphExec {
    hipGraphExec_t exec = nullptr;
    GraphExec() = default;
    GraphExec(const GraphExec &) = delete;
    GraphExec &operator=(const GraphExec &) = delete;
    ~GraphExec() {
        if (exec) (void)hipGraphExecDestroy(exec);
    }
};

// HIP events owned by a render: destroyed on every exit, errors included.
struct EventVec : std::vector<hipEvent_t> {
    explicit EventVec(size_t n = 0) : std::vector<hipEvent_t>(n, nullptr) {}
    EventVec(const EventVec &) = delete;
    EventVec &operator=(const EventVec &) = delete;
    ~EventVec() {
        for (hipEvent_t e : *this)
            if (e) (void)hipEventDestroy(e);
    }
};

// Frame lanes (cached and progressive renders): lane l runs on its own stream.
constexpr int kMaxLanes = 4;
hipStream_t lane_stream(const tpt_hip_scene *h, int l) {
    return l == 0 ? h->stream : l == 1 ? h->stream2 : h->streamx[l - 2];
}
void ensure_lane_streams(tpt_hip_scene *h, int lanes) {
    if (lanes > 1) ensure_stream2(h);
    for (int l = 2; l < lanes; l++)
        if (!h->streamx[l - 2]) HIP_OK(hipStreamCreateWithFlags(&h->streamx[l - 2], hipStreamNonBlocking));
}
// Path pool of lane l of the scene (lanes >= 1 add into pool 0's per-pixel sums).
PathBuffers *lane_pool(tpt_hip_scene *h, int l) { return l == 0 ? &h->pb : l == 1 ? &h->pb2 : &h->pbx[l - 2]; }
void sync_lanes(const tpt_hip_scene *h);
void ensure_lane_pool(tpt_hip_scene *h, int l, size_t slots, int pixels);
// Waits for everything queued on the scene's streams.
void sync_lanes(const tpt_hip_scene *h) {
    for (int l = 0; l < kMaxLanes; l++) {
        hipStream_t x = lane_stream(h, l);
        if (x) HIP_OK(hipStreamSynchronize(x));
    }
}

// A second path pool adds into the first pool's per-pixel sums.
void share_sums(PathBuffers &dst, const PathBuffers &src) {
    dst.P.sub_f = src.P.sub_f;
    dst.P.sub_d = src.P.sub_d;
    dst.P.pix_f = src.P.pix_f;
    dst.P.pix_d = src.P.pix_d;
}

// Second path pool and stream for two-stream renders: its own path state and
// queues, sharing pb's per-pixel accumulators (call after ensure_paths).
void ensure_paths2(tpt_hip_scene *h, size_t slots, int pixels) {
    ensure_stream2(h);
    PathBuffers &pb = h->pb2;
    if (!(pb.mem && pb.slots >= slots && pb.pixels == pixels)) {
        if (pb.mem) {
            HIP_OK(hipStreamSynchronize(h->stream2));
            HIP_OK(hipFree(pb.mem));
            pb.mem = nullptr;
        }
        alloc_paths(pb, slots, (size_t)pixels, false);
    }
    share_sums(pb, h->pb);
}

// Pool of lane l with `slots` paths for `pixels` pixels (call for lane 0 first).
void ensure_lane_pool(tpt_hip_scene *h, int l, size_t slots, int pixels) {
    if (l == 0) return ensure_paths(h, slots, pixels);
    if (l == 1) return ensure_paths2(h, slots, pixels);
    ensure_lane_streams(h, l + 1);
    PathBuffers &x = h->pbx[l - 2];
    if (!(x.mem && x.slots >= slots && x.pixels == pixels)) {
        sync_lanes(h);
        alloc_paths(x, slots, (size_t)pixels, false);
    }
    share_sums(x, h->pb);
}

int block_size_for(int spp) { // PixelSampler ctor (cpu/path_tracer.cpp:10-21)
    const int m = 16;
    if (spp < 2 * 2 * m) return 1;
    if (spp < 3 * 3 * m) return 2;
    if (spp < 4 * 4 * m) return 3;
    return 4;
}

// Sampler and integrator settings of one render (cpu/path_tracer.cpp:8-21, 102-104).
DRender make_render(const tpt_render_params *p, int spp, uint64_t seed, int pixels) {
    DRender R{};
    R.block = block_size_for(spp);
    R.nsub = spp / (R.block * R.block);
    R.spp_eff = R.nsub * R.block * R.block;
    R.seed = seed;
    R.max_depth = p->max_depth;
    R.rr_depth = p->rr_depth;
    R.shadow_bounded = p->shadow_bounded;
    R.output = p->output == TPT_OUT_HDR_MEAN ? 1 : 0;
    R.accum_fp64 = p->accum_fp64 ? 1 : 0;
    R.pixels = pixels;
    return R;
}

// Samples per pixel of one pipeline wave: as many as `cap` path slots hold.
// (Spreading the samples evenly over the waves instead, 1024x768 at 256 spp in
// 2^24 slots as 13 waves of 20 / 16 rather than 12 of 21 and one of 4, measured
// 605-615 against 614-625 Mpaths/s on three streams: the short last wave
// overlaps the others' tails.)
int wave_samples(int spp_eff, size_t cap, size_t pixels) {
    return (int)std::max<size_t>(1, std::min<size_t>((size_t)spp_eff, cap / pixels));
}

int grid_for(size_t n, int block, int cus = 256) {
    size_t g = (n + block - 1) / block;
    // one thread per item up to 65536 threads per CU (grid-stride beyond): the
    // hardware starts a new wave wherever one ends, which hides the shade
    // kernel's gather latency better than a few resident waves looping over
    // many items (room 1024x768x256: 2048 threads per CU 627, 8192 647,
    // uncapped 653 Mpaths/s)
    size_t cap = (size_t)cus * 65536 / block;
    if (g > cap) g = cap;
    if (g < 1) g = 1;
    return (int)g;
}

// Pipeline waves of one plain render (generate, max_depth + 1 x extend / shade /
// shadow, accumulate) over `npools` path pools on as many streams: wave w runs on
// pool (lane0 + w) % npools.  The waves share the per-pixel sums, so each wave's
// accumulate waits for the previous one's (`prev_acc`: index into `evs` of that
// accumulate's event, -1 before the first wave; updated) and the samples are added
// in order.  With `timing`, events bracket every extend / shade / shadow launch.
// A scene without a BVH (analytic primitives only) traces each wave's paths whole
// in k_megapath instead (same colours; `timing` brackets it as shading), unless
// traffic is counted or TPT_MEGAKERNEL=0.
struct PlainWaves {
    DPaths pools[kMaxLanes] = {};
    hipStream_t streams[kMaxLanes] = {};
    int npools = 1, block = 256, cus = 256;
    bool count = false, timing = false;
    bool megakernel = env_int("TPT_MEGAKERNEL", 1, 0, 1) != 0;
    // k_megapath with lane refill: waves per CU of its grid (0: one thread per path,
    // no refill) and the idle lanes of a wave that start their next paths together
    // (two-spheres box 512x384x64: one thread per path 4.65 ms; 32 waves per CU
    // refilling at 1 / 2 / 4 / 8 / 16 / 24 idle lanes 4.15 / 4.20 / 4.25-4.30 /
    // 4.28 / 4.38 / 4.55 ms; 16, 24, 48 and 64 waves per CU no better)
    int mega_waves = env_int("TPT_MEGA_WAVES", 32, 0, 1024);
    int mega_refill = env_int("TPT_MEGA_REFILL_LANES", 1, 1, 64);
    std::vector<std::pair<int, int>> ext_ev, shade_ev, shadow_ev;
    int waves = 0;
    static int record(EventVec &evs, hipStream_t s, bool timing) {
        hipEvent_t e;
        HIP_OK(hipEventCreateWithFlags(&e, timing ? hipEventDefault : hipEventDisableTiming));
        evs.push_back(e);
        HIP_OK(hipEventRecord(e, s));
        return (int)evs.size() - 1;
    }
    // returns the pool of the last wave
    int enqueue(const DScene &S, const DRender &R, int k, uint64_t base_mixed, int lane0, EventVec &evs, int &prev_acc) {
        const int pixels = R.pixels;
        int lane = lane0;
        for (int wave_base = 0, w = 0; wave_base < R.spp_eff; wave_base += k, w++) {
            lane = (lane0 + w) % npools;
            hipStream_t sw = streams[lane];
            const DPaths &P = pools[lane];
            auto ev = [&]() { return record(evs, sw, timing); };
            int k_active = std::min(k, R.spp_eff - wave_base);
            int n = pixels * k_active;
            int grid = grid_for((size_t)n, block, cus);
            if (megakernel && S.root == -1 && !count) {
                int e0 = timing ? ev() : -1;
                const bool refill = mega_waves > 0;
                const int mgrid = refill ? std::max(1, std::min(grid, cus * mega_waves * 64 / block)) : grid;
                launch_megapath(S, R, P, wave_base, n, base_mixed, mgrid, block, refill ? mega_refill : 0, sw);
                if (timing) shade_ev.push_back({e0, ev()});
            } else {
                launch_generate(S, R, P, wave_base, n, base_mixed, block, sw);
                int cur = 0;
                for (int b = 0; b <= R.max_depth; b++) {
                    int32_t *q = cur == 0 ? P.queue_a : P.queue_b;
                    int32_t *nx = cur == 0 ? P.queue_b : P.queue_a;
                    int e0 = timing ? ev() : -1;
                    launch_extend(S, P, q, cur, 1 - cur, grid, block, count, (uint64_t)n, sw);
                    int e1 = timing ? ev() : -1;
                    launch_shade(S, R, P, q, nx, cur, 1 - cur, grid, block, sw, nullptr);
                    int e2 = timing ? ev() : -1;
                    launch_shadow(S, P, grid, block, count, (uint64_t)n, sw);
                    int e3 = timing ? ev() : -1;
                    if (timing) {
                        ext_ev.push_back({e0, e1});
                        shade_ev.push_back({e1, e2});
                        shadow_ev.push_back({e2, e3});
                    }
                    cur = 1 - cur;
                }
            }
            if (npools > 1 && prev_acc >= 0) HIP_OK(hipStreamWaitEvent(sw, evs[prev_acc], 0));
            launch_accumulate(R, P, wave_base, k_active, w == 0 ? 1 : 0, sw);
            if (npools > 1) prev_acc = record(evs, sw, false);
            waves++;
        }
        return lane;
    }
};

} // namespace

extern "C" {

int tpt_hip_abi_version(void) { return TPT_ABI_VERSION; }

int tpt_hip_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

const char *tpt_hip_last_error(void) { return g_err.c_str(); }

int tpt_hip_scene_create(const tpt_scene_desc *desc, tpt_hip_scene **out) {
    if (!desc || !out) return set_err(TPT_ERR_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    tpt_hip_scene *h = nullptr;
    try {
        HostScene H;
        build_host_scene(*desc, H);
        int n = 0;
        if (hipGetDeviceCount(&n) != hipSuccess || n == 0)
            throw tpt_error(TPT_ERR_HIP, "hip is not available: no HIP device visible");
        h = new tpt_hip_scene();
        HIP_OK(hipGetDevice(&h->device));
        HIP_OK(hipDeviceGetAttribute(&h->cus, hipDeviceAttributeMultiprocessorCount, h->device));
        h->S.num_cu = h->cus;
        HIP_OK(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
        upload_scene(h, H);
        // keep the traversal kernels' dynamic LDS (node cache + stack) within one workgroup's limit
        int max_shared = 0;
        HIP_OK(hipDeviceGetAttribute(&max_shared, hipDeviceAttributeMaxSharedMemoryPerBlock, h->device));
        int stack_bytes = h->S.lds_stack ? kLdsRow * h->S.trav_block * 4 : 0;
        if (stack_bytes > max_shared) {
            h->S.lds_stack = 0;
            stack_bytes = 0;
        }
        h->S.lds_nodes = std::min(h->S.lds_nodes, (max_shared - stack_bytes) / 64);
        h->S.max_lds = max_shared; // one-ray-per-lane launches fit their LDS per block size
        // lanes (a path pool on its own stream each): plain renders rotate pipeline
        // waves, cached / progressive renders frames, over them, with the same image
        // as one lane (batches: up to three pools); TPT_STREAMS=1 keeps one.  Two
        // lanes against one took the room from 447 to 542 Mpaths/s when added.  Room 1024x768x256: 668 / 699 / 656 Mpaths/s with 2 / 3 / 4;
        // textured room 1920x1080, 16 / 64 cached frames: 54.7 / 216.9 ms with 2,
        // 50.0 / 191.9 with 3, 57.2 / 220.7 with 4
        h->streams = env_int("TPT_STREAMS", 3, 1, 4);
        HIP_OK(hipMalloc(&h->trace_counters, 64));
        *out = h;
        return TPT_OK;
    } catch (const tpt_error &e) {
        delete h;
        return set_err(e.code, e.what());
    } catch (const std::exception &e) {
        delete h;
        return set_err(TPT_ERR_RUNTIME, e.what());
    }
}

void tpt_hip_scene_destroy(tpt_hip_scene *scene) { delete scene; }

int tpt_hip_scene_info(const tpt_hip_scene *s, uint64_t *arena_bytes, uint64_t *num_nodes, uint64_t *num_triangles,
                       uint64_t *num_analytic, uint64_t *node_bytes) {
    if (!s) return set_err(TPT_ERR_INVALID_ARGUMENT, "null scene");
    if (arena_bytes) *arena_bytes = s->arena_bytes;
    if (num_nodes) *num_nodes = s->num_nodes;
    if (num_triangles) *num_triangles = s->num_tris;
    if (num_analytic) *num_analytic = s->num_analytic;
    if (node_bytes) *node_bytes = (uint64_t)kNodeFloat4[s->S.layout] * 16u;
    return TPT_OK;
}

int tpt_hip_render(tpt_hip_scene *h, const tpt_render_params *p, float *out_rgb, tpt_render_stats *stats) {
    if (!h || !p || !out_rgb) return set_err(TPT_ERR_INVALID_ARGUMENT, "null argument");
    if (p->spp <= 0) return set_err(TPT_ERR_INVALID_ARGUMENT, "spp must be positive");
    if (p->max_depth < 0 || p->max_depth > 1000) return set_err(TPT_ERR_INVALID_ARGUMENT, "max_depth out of range");
    try {
        auto t0 = std::chrono::steady_clock::now();
        HIP_OK(hipSetDevice(h->device));
        const DScene &S = h->S;
        int W = S.cam.width, Hh = S.cam.height;
        int pixels = W * Hh;
        DRender R = make_render(p, p->spp, p->seed, pixels);
        size_t cap = p->queue_capacity > 0 ? (size_t)p->queue_capacity : ((size_t)1 << 24);
        const int k = wave_samples(R.spp_eff, cap, (size_t)pixels);
        R.k = k;
        int block = p->block_size > 0 ? p->block_size : 256;
        if (block % 64 != 0 || block > 1024) return set_err(TPT_ERR_INVALID_ARGUMENT, "block_size must be a multiple of 64 <= 1024");
        ensure_paths(h, (size_t)pixels * k, pixels);
        // Several streams: pipeline wave w runs on lane w % npools (the scene's stream
        // and pool pb, then stream2 / pb2, ...), so one wave's latency-bound shading
        // overlaps another's VALU-bound traversal.  The waves share the per-pixel
        // sums; each accumulate waits for the previous wave's, keeping sample order.
        const int nwaves = (R.spp_eff + k - 1) / k;
        const int npools = std::max(1, std::min({h->streams, nwaves, kMaxLanes}));
        const bool two = npools > 1;
        for (int l = 1; l < npools; l++) ensure_lane_pool(h, l, (size_t)pixels * k, pixels);
        hipStream_t st = h->stream;
        uint64_t base_mixed = splitmix64(p->seed);
        bool timing = stats != nullptr;
        bool count = p->count_traffic != 0;
        if (count) HIP_OK(hipMemsetAsync(h->pb.P.counters + kCntTraffic, 0, 32, st));
        EventVec evs;
        auto ev_on = [&](hipStream_t s) {
            hipEvent_t e;
            HIP_OK(hipEventCreateWithFlags(&e, timing ? hipEventDefault : hipEventDisableTiming));
            evs.push_back(e);
            HIP_OK(hipEventRecord(e, s));
            return (int)evs.size() - 1;
        };
        if (two) {
            for (int l = 1; l < npools; l++)
                if (count) HIP_OK(hipMemsetAsync(lane_pool(h, l)->P.counters + kCntTraffic, 0, 32, st));
            const int e = ev_on(st);
            for (int l = 1; l < npools; l++) // the other pools start after this point of st
                HIP_OK(hipStreamWaitEvent(lane_stream(h, l), evs[e], 0));
        }
        PlainWaves pw;
        for (int l = 0; l < npools; l++) {
            pw.pools[l] = lane_pool(h, l)->P;
            pw.streams[l] = lane_stream(h, l);
        }
        pw.npools = npools;
        pw.block = block;
        pw.cus = h->cus;
        pw.count = count;
        pw.timing = timing;
        int prev_acc = -1;
        pw.enqueue(S, R, k, base_mixed, 0, evs, prev_acc);
        auto &ext_ev = pw.ext_ev, &shade_ev = pw.shade_ev, &shadow_ev = pw.shadow_ev;
        if (two) HIP_OK(hipStreamWaitEvent(st, evs[prev_acc], 0)); // the last accumulate may be on another stream
        launch_finalize(R, h->pb.P, h->pb.out, st);
        HIP_OK(hipGetLastError());
        HIP_OK(hipMemcpyAsync(out_rgb, h->pb.out, (size_t)pixels * 12, hipMemcpyDeviceToHost, st));
        HIP_OK(hipStreamSynchronize(st));
        if (two) sync_lanes(h);
        auto t1 = std::chrono::steady_clock::now();
        if (stats) {
            std::memset(stats, 0, sizeof(*stats));
            stats->render_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
            auto sum = [&](const std::vector<std::pair<int, int>> &v) {
                double s = 0;
                for (auto &pr : v) {
                    float ms = 0;
                    HIP_OK(hipEventElapsedTime(&ms, evs[pr.first], evs[pr.second]));
                    s += ms;
                }
                return s;
            };
            stats->extend_ms = sum(ext_ev);
            stats->shade_ms = sum(shade_ev);
            stats->shadow_ms = sum(shadow_ev);
            stats->paths = (uint64_t)pixels * (uint64_t)R.spp_eff;
            stats->extend_launches = ext_ev.size();
            if (count) {
                uint64_t sum4[4] = {0, 0, 0, 0};
                for (int l = 0; l < npools; l++) {
                    uint64_t c[4];
                    HIP_OK(hipMemcpy(c, lane_pool(h, l)->P.counters + kCntTraffic, 32, hipMemcpyDeviceToHost));
                    for (int i = 0; i < 4; i++) sum4[i] += c[i];
                }
                stats->nodes_visited = sum4[0];
                stats->tris_tested = sum4[1];
                stats->rays_closest = sum4[2];
                stats->rays_shadow = sum4[3];
            }
        }
        return TPT_OK;
    } catch (const tpt_error &e) {
        return set_err(e.code, e.what());
    } catch (const std::exception &e) {
        return set_err(TPT_ERR_RUNTIME, e.what());
    }
}

// ---------------------------------------------------------------- render without allocation
// One path pool of kNoAllocSlots paths and the per-pixel sums of kNoAllocPixels
// pixels in a fixed device array (`scratch`, one per device, held until the
// process ends), for callers that cannot allocate.  In host code `scratch` names
// a host-side stand-in, not the device array: the array's address on a device
// comes from hipGetSymbolAddress with that device current, once per device.
namespace {
constexpr size_t kNoAllocSlots = (size_t)1 << 19;
constexpr size_t kNoAllocPixels = (size_t)1 << 19;
constexpr PathLayout kNoAllocLayout = path_layout(kNoAllocSlots, kNoAllocPixels, false, false);
constexpr int kNoAllocMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kNoAllocMaxDevices] = {};
} // namespace

static __device__ __align__(256) uint8_t scratch[kNoAllocLayout.total];

namespace {
// Device address of `scratch` on `device` (current), looked up on the first call
// for that device.
uint8_t *scratch_on(int device) {
    if (device < 0 || device >= kNoAllocMaxDevices)
        throw tpt_error(TPT_ERR_UNSUPPORTED, "tpt_hip_render_noalloc: device index beyond " +
                                                 std::to_string(kNoAllocMaxDevices));
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void *ptr = nullptr;
        HIP_OK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        if (!ptr) throw tpt_error(TPT_ERR_HIP, "tpt_hip_render_noalloc: no device address for the scratch array");
        g_scratch_ptr[device] = ptr;
    }
    return (uint8_t *)g_scratch_ptr[device];
}
} // namespace

extern "C" int tpt_hip_render_noalloc(tpt_hip_scene *h, const tpt_render_params *p, float *out_rgb, void *stream) {
    if (!h || !p || !out_rgb) return set_err(TPT_ERR_INVALID_ARGUMENT, "null argument");
    if (p->spp <= 0) return set_err(TPT_ERR_INVALID_ARGUMENT, "spp must be positive");
    if (p->max_depth < 0 || p->max_depth > 1000) return set_err(TPT_ERR_INVALID_ARGUMENT, "max_depth out of range");
    if (p->count_traffic) return set_err(TPT_ERR_INVALID_ARGUMENT, "tpt_hip_render_noalloc: count_traffic is not supported");
    const int block = p->block_size > 0 ? p->block_size : 256;
    if (block % 64 != 0 || block > 1024) return set_err(TPT_ERR_INVALID_ARGUMENT, "block_size must be a multiple of 64 <= 1024");
    const DScene &S = h->S;
    const size_t pixels = (size_t)S.cam.width * (size_t)S.cam.height;
    if (pixels > kNoAllocPixels)
        return set_err(TPT_ERR_INVALID_ARGUMENT, "tpt_hip_render_noalloc: " + std::to_string(pixels) +
                                                     " pixels exceed its limit of " + std::to_string(kNoAllocPixels));
    if (p->queue_capacity < 0 || (size_t)p->queue_capacity > kNoAllocSlots)
        return set_err(TPT_ERR_INVALID_ARGUMENT, "tpt_hip_render_noalloc: queue_capacity must be in 0 .. " +
                                                     std::to_string(kNoAllocSlots));
    try {
        HIP_OK(hipSetDevice(h->device));
        uint8_t *base = scratch_on(h->device);
        hipStream_t st = (hipStream_t)stream;
        DPaths P{};
        bind_paths(base, kNoAllocLayout, false, false, P, nullptr);
        DRender R = make_render(p, p->spp, p->seed, (int)pixels);
        const size_t cap = p->queue_capacity > 0 ? (size_t)p->queue_capacity : kNoAllocSlots;
        // samples per pixel per pipeline wave; pixels * k <= max(cap, pixels) <= kNoAllocSlots
        const int k = (int)std::max<size_t>(1, std::min<size_t>((size_t)R.spp_eff, cap / pixels));
        R.k = k;
        // the array keeps the previous call's contents: the queue and refill counters start at zero
        launch_fill_u32(P.counters, kCntWords, 0u, st);
        const uint64_t base_mixed = splitmix64(p->seed);
        int first = 1;
        for (int wave_base = 0; wave_base < R.spp_eff; wave_base += k) {
            const int k_active = std::min(k, R.spp_eff - wave_base);
            const int n = (int)pixels * k_active;
            launch_generate(S, R, P, wave_base, n, base_mixed, block, st);
            const int grid = grid_for((size_t)n, block, h->cus);
            int cur = 0;
            for (int b = 0; b <= R.max_depth; b++) {
                int32_t *q = cur == 0 ? P.queue_a : P.queue_b;
                int32_t *nx = cur == 0 ? P.queue_b : P.queue_a;
                launch_extend(S, P, q, cur, 1 - cur, grid, block, false, (uint64_t)n, st);
                launch_shade(S, R, P, q, nx, cur, 1 - cur, grid, block, st, nullptr);
                launch_shadow(S, P, grid, block, false, (uint64_t)n, st);
                cur = 1 - cur;
            }
            launch_accumulate(R, P, wave_base, k_active, first, st);
            first = 0;
        }
        launch_finalize(R, P, out_rgb, st);
        HIP_OK(hipGetLastError());
        return TPT_OK;
    } catch (const tpt_error &e) {
        return set_err(e.code, e.what());
    } catch (const std::exception &e) {
        return set_err(TPT_ERR_RUNTIME, e.what());
    }
}

namespace {

// Device state of one batch render: scene / sampler tables, segment tables and
// path pools.  With several pools, pipeline waves rotate over them (pool 0 on
// the scene's stream, the others on streams of their own): one wave's shading
// overlaps the others' traversal.  All pools add into pool 0's per-pixel sums,
// each wave's accumulation after the previous wave's, so the images do not change.
// Pools of a batch at most (TPT_STREAMS permitting).  64 mixed rooms: 579 / 630
// Mpaths/s with 2 / 3 pools; 256 small rooms: 510 / 519.
#ifndef TPT_BATCH_POOLS
#define TPT_BATCH_POOLS 3
#endif
constexpr int kBatchPools = TPT_BATCH_POOLS;

// Batch path pool k of the handle `h` with at least `slots` paths and sums for
// `pixels` pixels; kept across batch calls, reallocated only when too small.
void ensure_batch_pool(tpt_hip_scene *h, int k, size_t slots, uint64_t pixels) {
    PathBuffers &x = h->batch_pb[k];
    if (x.mem && x.slots >= slots && (uint64_t)x.pixels >= pixels) return;
    if (x.mem) sync_lanes(h); // a previous batch may still run on the handle's streams
    alloc_paths(x, slots, (size_t)pixels, true);
}

struct BatchRun {
    static constexpr int kMaxPools = 4;
    void *meta = nullptr;                 // scene / sampler tables, pool 0's segments and wave start
    PathBuffers *pbs[kMaxPools] = {};     // path pools (owned by the first scene's handle); pool 0
                                          // holds the per-pixel sums the others add into
    DBatch B{};                           // pool 0's view of the batch
    hipStream_t st = nullptr;             // pool 0's stream
    void *xmeta[kMaxPools - 1] = {};      // pools 1..: segment table and wave start
    DBatch xB[kMaxPools - 1]{};
    hipStream_t xst[kMaxPools - 1] = {};
    uint64_t total_paths = 0, total_pixels = 0;
    std::vector<uint64_t> path_off, pix_off;
    int max_depth = 0, layout = 0, trav_block = 512, num_cu = 256, fp64 = 0;
    bool lds_stack = true; // first traversal-stack entries of each lane in LDS
    int mode = 4;          // traversal kernels (DScene::persistent)
    int chunk = 768;       // queue entries per wave of the chunked kernels
    int max_lds = 65536;
    int pools = 1;
    DBatch &batch(int k) { return k ? xB[k - 1] : B; }
    PathBuffers &pool(int k) { return *pbs[k]; }
    hipStream_t stream(int k) const { return k ? xst[k - 1] : st; }
    ~BatchRun() {
        for (int k = 0; k < pools; k++) // nothing queued may still use the buffers
            if (stream(k)) (void)hipStreamSynchronize(stream(k));
        if (meta) (void)hipFree(meta);
        for (int k = 0; k < kMaxPools - 1; k++)
            if (xmeta[k]) (void)hipFree(xmeta[k]);
    }
    // one more pool: its own path state and queues (`x`, adding into pool 0's
    // per-pixel sums), segment table and wave start
    void add_pool(PathBuffers *x, hipStream_t stream) {
        const int k = pools;
        pbs[k] = x;
        share_sums(*x, *pbs[0]);
        size_t seg_bytes = align256(4 * (size_t)kSegFields * (size_t)B.seg_stride);
        HIP_OK(hipMalloc(&xmeta[k - 1], seg_bytes + 16));
        HIP_OK(hipMemset(xmeta[k - 1], 0, seg_bytes + 16));
        DBatch &X = xB[k - 1];
        X = B;
        X.seg = (int32_t *)xmeta[k - 1];
        X.state = (uint64_t *)((uint8_t *)xmeta[k - 1] + seg_bytes);
        xst[k - 1] = stream;
        pools = k + 1;
        for (int q = 0; q < pools; q++) batch(q).pools = pools;
    }
    int block = 256; // generate / shade workgroup size of the pipeline waves
    // the pipeline part of one wave (every bounce's kernels) on stream s
    void launch_wave(int k, hipStream_t s) const {
        launch_batch_wave(xB_or_B(k), pbs[k]->P, max_depth, layout, block, trav_block, lds_stack, mode, chunk, max_lds,
                          num_cu, total_pixels, s);
    }
    const DBatch &xB_or_B(int k) const { return k ? xB[k - 1] : B; }
    // waves of capacity `cap` with workgroup size `blk` (sets every pool's view)
    void shape(int cap, int blk) {
        block = blk;
        for (int k = 0; k < pools; k++) batch(k).cap = cap;
    }
    // capture the pipeline part of one wave of (cap, block), per pool, into executable graphs
    void capture(int cap, int block, hipGraphExec_t ge[kMaxPools]) {
        shape(cap, block);
        for (int k = 0; k < pools; k++) {
            hipGraph_t g = nullptr;
            HIP_OK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            launch_wave(k, st);
            HIP_OK(hipStreamEndCapture(st, &g));
            hipError_t e = hipGraphInstantiate(&ge[k], g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            HIP_OK(e);
        }
    }
    // restart at path `start` with zeroed accumulators: pool k's first wave starts
    // at start + k * cap (the streams are drained first: the host copies are not
    // ordered with the non-blocking streams)
    void reset(int cap, uint64_t start = 0) {
        for (int k = 0; k < pools; k++) HIP_OK(hipStreamSynchronize(stream(k)));
        for (int k = 0; k < pools; k++) {
            uint64_t s0[2] = {start + (uint64_t)k * (uint64_t)cap, total_paths};
            HIP_OK(hipMemcpy(batch(k).state, s0, sizeof(s0), hipMemcpyHostToDevice));
        }
        launch_batch_zero(B, pbs[0]->P, fp64, total_pixels, st);
    }
    int waves(int cap) const { return (int)((total_paths + (uint64_t)cap - 1) / (uint64_t)cap); }
    // the reset of reset() as kernels (for a captured graph): pool k's first wave
    // starts at start + k * cap, accumulators zeroed
    void reset_on_device(int cap, uint64_t start = 0) {
        for (int k = 0; k < pools; k++) launch_batch_state(batch(k), start + (uint64_t)k * (uint64_t)cap, total_paths, st);
        launch_batch_zero(B, pbs[0]->P, fp64, total_pixels, st);
    }
    // enqueue waves [0, W): wave w's pipeline (its per-pool graph ge[k], or the
    // kernels themselves when ge is null, e.g. while st is being captured) on pool
    // w % pools's stream, then its tail after the previous wave's tail; stamps[w]
    // (if given) gets the device clock after wave w's tail
    void run(hipGraphExec_t *ge, int W, uint64_t *stamps) {
        std::vector<hipEvent_t> tail(W);
        for (auto &e : tail) HIP_OK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        if (pools > 1) { // the other pools start after the reset on st
            hipEvent_t e0;
            HIP_OK(hipEventCreateWithFlags(&e0, hipEventDisableTiming));
            HIP_OK(hipEventRecord(e0, st));
            for (int k = 1; k < pools; k++) HIP_OK(hipStreamWaitEvent(stream(k), e0, 0));
            HIP_OK(hipEventDestroy(e0));
        }
        for (int w = 0; w < W; w++) {
            const int k = w % pools;
            hipStream_t s = stream(k);
            if (ge) HIP_OK(hipGraphLaunch(ge[k], s));
            else launch_wave(k, s);
            if (pools > 1 && w > 0) HIP_OK(hipStreamWaitEvent(s, tail[w - 1], 0));
            launch_batch_tail(batch(k), pool(k).P, total_pixels, s);
            if (stamps) launch_stamp(stamps + w, s);
            HIP_OK(hipEventRecord(tail[w], s));
        }
        if (pools > 1 && W > 0) HIP_OK(hipStreamWaitEvent(st, tail[W - 1], 0)); // st continues after the last wave
        for (auto &e : tail) HIP_OK(hipEventDestroy(e));
    }
};

// Paths per ms of (cap, block) over the paths [start, start + span): every
// candidate is timed on the same paths (span is a multiple of every candidate
// capacity, or reaches the batch's end), so the comparison does not depend on
// which scenes a candidate's waves happen to cover.
double tune_trial(BatchRun &run, int cap, int block, uint64_t start, uint64_t span) {
    hipGraphExec_t ge[BatchRun::kMaxPools] = {};
    run.capture(cap, block, ge);
    const uint64_t end = std::min(run.total_paths, start + span);
    const int reps = (int)((end - start + (uint64_t)cap - 1) / (uint64_t)cap);
    run.reset(cap, start);
    hipEvent_t e0, e1;
    HIP_OK(hipEventCreate(&e0));
    HIP_OK(hipEventCreate(&e1));
    HIP_OK(hipEventRecord(e0, run.st));
    run.run(ge, reps, nullptr);
    HIP_OK(hipEventRecord(e1, run.st));
    HIP_OK(hipEventSynchronize(e1));
    float ms = 0;
    HIP_OK(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    for (int k = 0; k < run.pools; k++) (void)hipGraphExecDestroy(ge[k]);
    uint64_t paths = std::min<uint64_t>((uint64_t)reps * (uint64_t)cap, run.total_paths - start);
    return ms > 0 ? (double)paths / (double)ms : 0.0;
}

} // namespace

int tpt_hip_render_batch(tpt_hip_scene *const *scenes_in, const tpt_batch_item *items_in, int32_t n,
                         const tpt_render_params *p, float *const *out_in, tpt_batch_stats *stats,
                         double *scene_ms) {
    if (!scenes_in || !items_in || !p || !out_in || n <= 0) return set_err(TPT_ERR_INVALID_ARGUMENT, "null argument");
    for (int i = 0; i < n; i++)
        if (!scenes_in[i]) return set_err(TPT_ERR_INVALID_ARGUMENT, "null scene or output");
    // Paths run shortest scene first (fewest paths), which minimises the mean
    // completion time; the images do not depend on the order.
    std::vector<int> order(n);
    std::iota(order.begin(), order.end(), 0);
    auto cost = [&](int i) {
        const DScene &S = scenes_in[i]->S;
        int b = block_size_for(std::max(1, items_in[i].spp));
        return (uint64_t)S.cam.width * (uint64_t)S.cam.height * (uint64_t)(items_in[i].spp / (b * b) * b * b);
    };
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return cost(a) < cost(b); });
    std::vector<tpt_hip_scene *> scenes(n);
    std::vector<tpt_batch_item> items(n);
    std::vector<float *> out_rgb(n);
    for (int k = 0; k < n; k++) {
        scenes[k] = scenes_in[order[k]];
        items[k] = items_in[order[k]];
        out_rgb[k] = out_in[order[k]];
    }
    if (n > 1024) return set_err(TPT_ERR_INVALID_ARGUMENT, "at most 1024 scenes per batch");
    if (p->max_depth < 0 || p->max_depth > 1000) return set_err(TPT_ERR_INVALID_ARGUMENT, "max_depth out of range");
    if (p->block_size != 0 && (p->block_size % 64 != 0 || p->block_size > 1024))
        return set_err(TPT_ERR_INVALID_ARGUMENT, "block_size must be a multiple of 64 <= 1024");
    if (p->queue_capacity < 0) return set_err(TPT_ERR_INVALID_ARGUMENT, "queue_capacity must not be negative");
    for (int i = 0; i < n; i++) {
        if (!scenes[i] || !out_rgb[i]) return set_err(TPT_ERR_INVALID_ARGUMENT, "null scene or output");
        if (items[i].spp <= 0) return set_err(TPT_ERR_INVALID_ARGUMENT, "spp must be positive");
        if (scenes[i]->device != scenes[0]->device) return set_err(TPT_ERR_INVALID_ARGUMENT, "scenes on different devices");
        if (scenes[i]->S.layout != scenes[0]->S.layout)
            return set_err(TPT_ERR_INVALID_ARGUMENT, "scenes with different BVH node layouts");
    }
    try {
        auto t0 = std::chrono::steady_clock::now();
        HIP_OK(hipSetDevice(scenes[0]->device));
        BatchRun run;
        run.st = scenes[0]->stream;
        run.max_depth = p->max_depth;
        run.layout = scenes[0]->S.layout;
        run.trav_block = scenes[0]->S.trav_block;
        run.lds_stack = scenes[0]->S.lds_stack != 0;
        run.mode = scenes[0]->S.persistent;
        run.chunk = scenes[0]->S.chunk;
        run.max_lds = scenes[0]->S.max_lds;
        run.num_cu = scenes[0]->cus;
        run.fp64 = p->accum_fp64 ? 1 : 0;
        std::vector<DScene> table(n);
        std::vector<DRender> rend(n);
        std::vector<uint64_t> seeds(n);
        run.path_off.assign(n + 1, 0);
        run.pix_off.assign(n + 1, 0);
        // the device table stores every array as a byte offset from the lowest arena
        uintptr_t base = (uintptr_t)scenes[0]->arena;
        for (int i = 1; i < n; i++) base = std::min(base, (uintptr_t)scenes[i]->arena);
        auto to_offset = [&](auto &ptr) { ptr = (std::remove_reference_t<decltype(ptr)>)((uintptr_t)ptr - base); };
        for (int i = 0; i < n; i++) {
            table[i] = scenes[i]->S;
            DScene &T = table[i];
            to_offset(T.nodes), to_offset(T.tris_bvh), to_offset(T.tris), to_offset(T.tri_info), to_offset(T.shade_rec);
            to_offset(T.normals), to_offset(T.texcoords), to_offset(T.gid_local), to_offset(T.analytic);
            to_offset(T.materials), to_offset(T.images), to_offset(T.texels), to_offset(T.delta);
            to_offset(T.lights), to_offset(T.env_marg), to_offset(T.env_cond), to_offset(T.env_gmarg);
            to_offset(T.env_gcond), to_offset(T.light_shapes);
            table[i].lds_nodes = 0; // no LDS node cache: wavefronts of one workgroup trace different scenes
            int px = table[i].cam.width * table[i].cam.height;
            rend[i] = make_render(p, items[i].spp, items[i].seed, px);
            seeds[i] = splitmix64(items[i].seed);
            run.path_off[i + 1] = run.path_off[i] + (uint64_t)px * (uint64_t)rend[i].spp_eff;
            run.pix_off[i + 1] = run.pix_off[i] + (uint64_t)px;
        }
        // Scenes of at least 2^TPT_BATCH_LARGE paths (default 2^23; they sort last)
        // fill the GPU on their own: they run as plain pipeline waves of their own
        // (the single-scene kernels, no segment tables) on the batch's pools after
        // the batched scenes; the smaller scenes [0, nb) share batched waves.
        const uint64_t large = (uint64_t)1 << env_int("TPT_BATCH_LARGE", 23, 10, 40);
        int nb = n;
        while (nb > 0 && run.path_off[nb] - run.path_off[nb - 1] >= large) nb--;
        const uint64_t all_pixels = run.pix_off[n];
        run.total_paths = run.path_off[nb];
        run.total_pixels = run.pix_off[nb];
        // plain waves: as tpt_hip_render's default, up to 2^24 path slots
        constexpr size_t kPlainCap = (size_t)1 << 24;
        size_t plain_slots = 0;
        for (int i = nb; i < n; i++) {
            const size_t px = (size_t)rend[i].pixels;
            rend[i].k = wave_samples(rend[i].spp_eff, kPlainCap, px);
            plain_slots = std::max(plain_slots, px * (size_t)rend[i].k);
        }
        // tables in one allocation
        size_t off = 0;
        auto take = [&](size_t bytes) {
            size_t o = off;
            off = align256(off + bytes);
            return o;
        };
        size_t o_sc = take(sizeof(DScene) * n), o_r = take(sizeof(DRender) * n), o_po = take(8 * (n + 1)),
               o_xo = take(8 * (n + 1)), o_sd = take(8 * n), o_sg = take(4 * (size_t)kSegFields * seg_stride_for(n)), o_st = take(16);
        HIP_OK(hipMalloc(&run.meta, off));
        uint8_t *m = (uint8_t *)run.meta;
        HIP_OK(hipMemcpy(m + o_sc, table.data(), sizeof(DScene) * n, hipMemcpyHostToDevice));
        HIP_OK(hipMemcpy(m + o_r, rend.data(), sizeof(DRender) * n, hipMemcpyHostToDevice));
        HIP_OK(hipMemcpy(m + o_po, run.path_off.data(), 8 * (n + 1), hipMemcpyHostToDevice));
        HIP_OK(hipMemcpy(m + o_xo, run.pix_off.data(), 8 * (n + 1), hipMemcpyHostToDevice));
        HIP_OK(hipMemcpy(m + o_sd, seeds.data(), 8 * n, hipMemcpyHostToDevice));
        HIP_OK(hipMemsetAsync(m + o_sg, 0, 4 * (size_t)kSegFields * seg_stride_for(n), run.st));
        DBatch &B = run.B;
        B.scenes = (const DScene *)(m + o_sc);
        B.base = (const uint8_t *)base;
        B.renders = (const DRender *)(m + o_r);
        B.path_off = (const uint64_t *)(m + o_po);
        B.pix_off = (const uint64_t *)(m + o_xo);
        B.seed_mixed = (const uint64_t *)(m + o_sd);
        B.seg = (int32_t *)(m + o_sg);
        B.state = (uint64_t *)(m + o_st);
        B.n = nb;
        B.seg_stride = seg_stride_for(n);
        B.pools = 1;
        // candidate capacities: 2^18, 2^20, 2^21, 2^22 path slots (the batch
        // configuration's range; TPT_BATCH_MAX_CAP raises the top to 2^23 .. 2^26),
        // never (much) more than the batch needs
        auto round_cap = [&](uint64_t c) {
            return (int)std::max<uint64_t>(256, std::min<uint64_t>(c, (run.total_paths + 255) / 256 * 256));
        };
        std::vector<int> caps;
        if (p->queue_capacity > 0) caps.push_back(round_cap((uint64_t)p->queue_capacity));
        else {
            const int top = env_int("TPT_BATCH_MAX_CAP", 22, 20, 26);
            for (int e : {18, 20, 21, 22, 23, 24, 25, 26}) {
                if (e > top) break;
                int c = round_cap(1ull << e);
                if (caps.empty() || caps.back() != c) caps.push_back(c);
            }
        }
        std::vector<int> blocks;
        if (p->block_size > 0) blocks.push_back(p->block_size);
        else blocks = {64, 128, 256, 512};
        // several waves in flight: one pool per lane of the first scene, kept on its
        // handle across calls (reallocated only when a batch needs more)
        const size_t max_cap = std::max((size_t)*std::max_element(caps.begin(), caps.end()), plain_slots);
        const int npools = run.total_paths > (uint64_t)caps[0] || nb < n ? std::min(scenes[0]->streams, kBatchPools) : 1;
        ensure_lane_streams(scenes[0], npools);
        for (int k = 0; k < npools; k++) ensure_batch_pool(scenes[0], k, max_cap, all_pixels);
        run.pbs[0] = &scenes[0]->batch_pb[0];
        for (int k = 1; k < npools; k++) run.add_pool(&scenes[0]->batch_pb[k], lane_stream(scenes[0], k));
        int best_cap = caps[0], best_block = blocks.size() == 1 ? blocks[0] : 256;
        double tune_ms = 0;
        if (nb > 0 && (caps.size() > 1 || blocks.size() > 1)) {
            auto tt = std::chrono::steady_clock::now();
            double best = -1;
            // every trial times the same 2^27 paths around the batch's middle (the
            // whole batch when it is smaller): scenes run shortest first, so the
            // first paths belong to the small scenes
            const uint64_t span = std::min<uint64_t>(run.total_paths, (uint64_t)1 << 27);
            const uint64_t start = (run.total_paths - span) / 2;
            for (int c : caps) { // capacity first, at the middle block size
                double r = tune_trial(run, c, best_block, start, span);
                if (r > best) {
                    best = r;
                    best_cap = c;
                }
            }
            if (blocks.size() > 1) {
                for (int b : blocks) {
                    if (b == 256) continue; // measured above
                    double r = tune_trial(run, best_cap, b, start, span);
                    if (r > best) {
                        best = r;
                        best_block = b;
                    }
                }
            }
            tune_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - tt).count();
        }
        // The render: the batched scenes' pipeline waves (one captured wave per
        // pool, replayed from the host), then the large scenes' plain waves, over
        // the pools.  TPT_BATCH_GRAPH=1 captures the whole render from the pools'
        // streams into ONE hipGraph and launches it once instead: the images are
        // the same, but ROCm's graph executor overlaps the pools' kernels less
        // than the streams do (64 mixed rooms 3,141 against 3,107 ms; the city at
        // 256 spp as a batch of one 1,413 against 1,055 ms).  Completion times
        // come from device clock stamps: [0] the start, [1 + w] after batched
        // wave w, [1 + W + i] after large scene i.
        auto tr = std::chrono::steady_clock::now();
        const bool one_graph = env_int("TPT_BATCH_GRAPH", 0, 0, 1) != 0;
        int plain_waves = 0;
        hipGraphExec_t ge[BatchRun::kMaxPools] = {};
        const int W = nb > 0 ? run.waves(best_cap) : 0;
        DeviceBuffer stamp_buf((size_t)(2 + W + n) * 8);
        uint64_t *stamps = (uint64_t *)stamp_buf.ptr;
        if (nb > 0 && !one_graph) {
            run.capture(best_cap, best_block, ge);
            run.reset(best_cap);
        }
        hipGraph_t graph = nullptr;
        GraphExec whole;
        if (one_graph) HIP_OK(hipStreamBeginCapture(run.st, hipStreamCaptureModeThreadLocal));
        try {
            launch_stamp(stamps, run.st);
            if (nb > 0) {
                if (one_graph) {
                    run.shape(best_cap, best_block);
                    run.reset_on_device(best_cap);
                }
                run.run(one_graph ? nullptr : ge, W, stamps + 1);
                launch_batch_finalize(B, run.pbs[0]->P, run.pbs[0]->out, run.total_pixels, run.st);
            }
            // large scenes: plain pipeline waves on the same pools, continuing the
            // rotation; each scene's sums and image are its range of the pools' arrays
            EventVec pev;
            if (nb < n) {
                PlainWaves pw;
                pw.npools = run.pools;
                pw.block = 256;
                pw.cus = run.num_cu;
                for (int k = 0; k < run.pools; k++) pw.streams[k] = run.stream(k);
                if (run.pools > 1) { // every pool's plain waves follow what is queued on st so far
                    const int e = PlainWaves::record(pev, run.st, false);
                    for (int k = 1; k < run.pools; k++) HIP_OK(hipStreamWaitEvent(run.stream(k), pev[e], 0));
                }
                int lane0 = W % run.pools;
                for (int i = nb; i < n; i++) {
                    const size_t o3 = 3 * run.pix_off[i];
                    for (int k = 0; k < run.pools; k++) {
                        DPaths P = run.pool(k).P;
                        P.sub_f += o3, P.pix_f += o3, P.sub_d += o3, P.pix_d += o3;
                        pw.pools[k] = P;
                    }
                    int prev = -1;
                    const int last = pw.enqueue(scenes[i]->S, rend[i], rend[i].k, seeds[i], lane0, pev, prev);
                    // the last accumulate runs on the last wave's stream after all earlier ones
                    launch_finalize(rend[i], pw.pools[last], run.pbs[0]->out + o3, run.stream(last));
                    launch_stamp(stamps + 1 + W + i, run.stream(last));
                    lane0 = (last + 1) % run.pools;
                }
                plain_waves = pw.waves;
                for (int k = 0; k < run.pools; k++) { // st continues after every pool's last plain wave
                    if (run.stream(k) == run.st) continue;
                    const int e = PlainWaves::record(pev, run.stream(k), false);
                    HIP_OK(hipStreamWaitEvent(run.st, pev[e], 0));
                }
            }
            HIP_OK(hipGetLastError());
        } catch (...) {
            if (one_graph) { // end the capture before the error leaves
                hipGraph_t g = nullptr;
                (void)hipStreamEndCapture(run.st, &g);
                if (g) (void)hipGraphDestroy(g);
                (void)hipGetLastError();
            }
            throw;
        }
        if (one_graph) {
            HIP_OK(hipStreamEndCapture(run.st, &graph));
            hipError_t e = hipGraphInstantiate(&whole.exec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            HIP_OK(e);
            HIP_OK(hipGraphLaunch(whole.exec, run.st));
        }
        for (int i = 0; i < n; i++)
            HIP_OK(hipMemcpyAsync(out_rgb[i], run.pbs[0]->out + 3 * run.pix_off[i], (run.pix_off[i + 1] - run.pix_off[i]) * 12,
                                  hipMemcpyDeviceToHost, run.st));
        HIP_OK(hipStreamSynchronize(run.st));
        auto t1 = std::chrono::steady_clock::now();
        if (scene_ms) {
            std::vector<uint64_t> ts((size_t)(1 + W + n));
            HIP_OK(hipMemcpy(ts.data(), stamps, ts.size() * 8, hipMemcpyDeviceToHost));
            int khz = 100000;
            if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, scenes[0]->device) != hipSuccess || khz <= 0)
                khz = 100000;
            for (int i = 0; i < n; i++) {
                const uint64_t t = i < nb ? ts[1 + (run.path_off[i + 1] - 1) / (uint64_t)best_cap] : ts[1 + W + i];
                scene_ms[order[i]] = (double)(t - ts[0]) / (double)khz;
            }
        }
        for (int k = 0; k < run.pools; k++) // none captured with one graph or no batched scene
            if (ge[k]) (void)hipGraphExecDestroy(ge[k]);
        if (stats) {
            std::memset(stats, 0, sizeof(*stats));
            stats->render_ms = std::chrono::duration<double, std::milli>(t1 - tr).count();
            stats->tune_ms = tune_ms;
            stats->paths = run.path_off[n];
            stats->queue_capacity = best_cap;
            stats->block_size = best_block;
            stats->waves = W + plain_waves;
            stats->graphs = one_graph ? 1 : 0;
        }
        (void)t0;
        return TPT_OK;
    } catch (const tpt_error &e) {
        return set_err(e.code, e.what());
    } catch (const std::exception &e) {
        return set_err(TPT_ERR_RUNTIME, e.what());
    }
}

namespace {

// Network storage of a scene handle (allocated once).
void cache_alloc_weights(tpt_hip_scene *h) {
    CacheState &c = h->cache;
    if (c.wmem) return;
    size_t off = 0;
    auto take = [&](size_t bytes) {
        size_t o = off;
        off = align256(off + bytes);
        return o;
    };
    size_t o_w = take(4 * kNrcParams), o_m = take(4 * kNrcParams), o_v = take(4 * kNrcParams),
           o_g = take(4 * kNrcParams), o_wt = take(2 * kNrcParams), o_wb = take(2 * kNrcParams), o_l = take(16),
           o_p = take(4 * (size_t)kNrcParams * kNrcMaxTrainGroups);
    HIP_OK(hipMalloc(&c.wmem, off));
    HIP_OK(hipMemset(c.wmem, 0, off));
    uint8_t *b = (uint8_t *)c.wmem;
    c.W.w = (float *)(b + o_w);
    c.W.m = (float *)(b + o_m);
    c.W.v = (float *)(b + o_v);
    c.W.grad = (float *)(b + o_g);
    c.W.wt = (uint16_t *)(b + o_wt);
    c.W.wb = (uint16_t *)(b + o_wb);
    c.W.loss = (float *)(b + o_l);
    c.W.partial = (float *)(b + o_p);
}

// Uploads FP32 weights [l][in][out] with their BF16 copies and clears the Adam state.
void cache_set_weights(tpt_hip_scene *h, const float *w) {
    CacheState &c = h->cache;
    std::vector<float> wf(w, w + kNrcParams);
    std::vector<uint16_t> wt(kNrcParams), wb(kNrcParams);
    for (int ly = 0; ly < 5; ly++)
        for (int in = 0; in < 64; in++)
            for (int out = 0; out < 64; out++) {
                int i = ly * 4096 + in * 64 + out;
                if (ly == 4 && out >= 3) wf[i] = 0.f;
                uint16_t b = f32_to_bf16_bits(wf[i]);
                wb[i] = b;
                wt[ly * 4096 + out * 64 + in] = b;
            }
    sync_lanes(h); // frames on any lane may still train the network
    HIP_OK(hipMemcpy(c.W.w, wf.data(), 4 * kNrcParams, hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(c.W.wt, wt.data(), 2 * kNrcParams, hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(c.W.wb, wb.data(), 2 * kNrcParams, hipMemcpyHostToDevice));
    std::vector<float> zero(kNrcParams, 0.f);
    HIP_OK(hipMemcpy(c.W.m, zero.data(), 4 * kNrcParams, hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(c.W.v, zero.data(), 4 * kNrcParams, hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(c.W.grad, zero.data(), 4 * kNrcParams, hipMemcpyHostToDevice));
    c.step = 0;
}

// Hidden layers: He-uniform U(-sqrt(6 / 64), sqrt(6 / 64)), seeded; the output
// layer starts at zero, so an untrained cache predicts no radiance.
void cache_init_weights(tpt_hip_scene *h, uint64_t seed) {
    std::vector<float> w(kNrcParams, 0.f);
    uint64_t s = seed;
    const float a = std::sqrt(6.f / 64.f);
    for (int i = 0; i < 4 * 4096; i++) {
        s = splitmix64(s + 0x9E3779B97F4A7C15ull);
        float u = (float)((s >> 40) * (1.0 / 16777216.0)); // [0, 1)
        w[i] = (2.f * u - 1.f) * a;
    }
    cache_set_weights(h, w.data());
}

// Staging buffer set `lane` (one per path pool) for `slots` paths and
// `train_cap` training records.
void cache_alloc_buffers(tpt_hip_scene *h, size_t slots, int train_cap, int lane) {
    CacheBufs &c = h->cache.buf[lane];
    if (c.bmem && c.slots >= slots && c.train_cap >= train_cap) return;
    if (c.bmem) {
        sync_lanes(h);
        HIP_OK(hipFree(c.bmem));
        c.bmem = nullptr;
    }
    size_t off = 0;
    auto take = [&](size_t bytes) {
        size_t o = off;
        off = align256(off + bytes);
        return o;
    };
    size_t o_f = take(slots * kNrcInputs * 2), o_ft = take(slots * kNrcInputs * 2), o_T = take(slots * 16),
           o_C = take(slots * 16), o_p = take(slots * 16),
           o_q = take(slots * 4), o_t = take((size_t)train_cap * 4), o_tg = take((size_t)train_cap * 16),
           o_c = take(16);
    HIP_OK(hipMalloc(&c.bmem, off));
    uint8_t *b = (uint8_t *)c.bmem;
    HIP_OK(hipMemset(b + o_c, 0, 16));
    c.feat = (uint16_t *)(b + o_f);
    c.feat_t = (uint16_t *)(b + o_ft);
    c.rec_T = (float4 *)(b + o_T);
    c.rec_C = (float4 *)(b + o_C);
    c.pred = (float4 *)(b + o_p);
    c.q_slot = (int32_t *)(b + o_q);
    c.t_slot = (int32_t *)(b + o_t);
    c.target = (float4 *)(b + o_tg);
    c.counters = (uint32_t *)(b + o_c);
    c.slots = slots;
    c.train_cap = train_cap;
}

} // namespace

namespace {

// Progressive frames of one sample per pixel on a scene: frame f is sample f of
// the plain sampler with one sub-pixel block (b = 1), so its camera ray and
// stream do not depend on how many frames are rendered in total.
struct FrameSetup {
    DRender R{};
    DCache C{};
    tpt_cache_params cp{};
    bool use_cache = false;
    int pixels = 0, fp64 = 0;
    int group = 1;    // frames traced together in one pipeline wave (group * pixels paths per launch)
    size_t slots = 0; // group * pixels: path slots of a lane's pool
    int lanes = 1;    // frame groups rotate over the lanes: path pools, streams and cache staging buffers
    uint64_t base_mixed = 0;
    int last_bounce = 0; // last bounce a path can reach (later bounces would trace empty queues)
};
constexpr int kFrameBlock = 256;
constexpr int kMaxFrameGroup = 4;

// Orders frames that run on different streams: `acc` is recorded after the
// latest frame's accumulation (the per-pixel sums are shared and added in frame
// order), `start` where the other lanes' streams join the scene's stream.  The
// training order is the scene's (tpt_hip_scene::train_ev).
struct FrameChain {
    hipEvent_t acc = nullptr, start = nullptr;
    bool any = false; // a frame has been enqueued: later frames wait on the events
    int next_lane = 0, last_lane = 0; // lane of the next frame group, of the latest one
    int last_frames = 0;              // frames of the latest group
    FrameChain() = default;
    FrameChain(const FrameChain &) = delete;
    FrameChain &operator=(const FrameChain &) = delete;
    void create() { // each missing event (a failed call leaves the others to the destructor)
        for (hipEvent_t *e : {&acc, &start})
            if (!*e) HIP_OK(hipEventCreateWithFlags(e, hipEventDisableTiming));
    }
    ~FrameChain() {
        for (hipEvent_t e : {acc, start})
            if (e) (void)hipEventDestroy(e);
    }
};

// Validates the parameters (cache may be null: no cache; its `frames` field is
// not used) and prepares the scene's network and cache buffers.
void setup_frames(tpt_hip_scene *h, const tpt_render_params *p, const tpt_cache_params *cp, FrameSetup &fs) {
    auto bad = [](const char *m) { throw tpt_error(TPT_ERR_INVALID_ARGUMENT, m); };
    if (p->max_depth < 0 || p->max_depth > 1000) bad("max_depth out of range");
    fs = FrameSetup{};
    fs.use_cache = cp && cp->cache_depth >= 0;
    if (cp) {
        if (cp->train_records <= 0 || cp->train_records > (1 << 20)) bad("train_records must be in 1 .. 2^20");
        if (!(cp->learning_rate >= 0.f)) bad("learning_rate must not be negative");
        if (cp->train_steps < 1 || cp->train_steps > cp->train_records)
            bad("train_steps must be in 1 .. train_records");
        if (cp->suffix_len < 0) bad("suffix_len must not be negative");
        fs.cp = *cp;
    }
    HIP_OK(hipSetDevice(h->device));
    const DScene &S = h->S;
    fs.pixels = S.cam.width * S.cam.height;
    fs.R = make_render(p, 1, p->seed, fs.pixels);
    fs.R.block = 1;
    fs.R.nsub = INT32_MAX; // one sub-pixel: sample s never leaves it
    fs.R.spp_eff = 1;
    fs.R.output = 1;
    fs.R.k = 1;
    // Frames per pipeline wave: one frame of a 1920x1080 image is 2M paths, and
    // launches that small lose to their fill and drain, so plain frames are
    // traced in groups of up to about 4M paths (at most kMaxFrameGroup;
    // TPT_FRAME_GROUP overrides).  A group's frames are the same samples as one
    // by one and accumulate in frame order: without the cache the image does not
    // depend on the grouping.  Textured room 1920x1080, 64 plain frames on three
    // streams: groups of 1 / 2 / 4 take 341 / 323 / 317 ms.  With the cache a
    // group's frames share the weights of the training before the group, and
    // groups do not pay (256 cached frames: 675 / 672 / 693 ms, the frames bound
    // by their three bounces' path kernels), so cached frames go one by one.
    const int gauto = fs.use_cache ? 1
                                   : (int)std::min<size_t>(kMaxFrameGroup,
                                                           std::max<size_t>(1, ((size_t)1 << 22) / fs.pixels));
    fs.group = env_int("TPT_FRAME_GROUP", gauto, 1, kMaxFrameGroup);
    fs.slots = (size_t)fs.group * fs.pixels;
    fs.base_mixed = splitmix64(p->seed);
    fs.fp64 = p->accum_fp64 ? 1 : 0;
    fs.lanes = std::min(h->streams, kMaxLanes);
    ensure_lane_streams(h, fs.lanes);
    if (fs.use_cache) {
        cache_alloc_weights(h);
        if (cp->reset || h->cache.step == 0) cache_init_weights(h, cp->weight_seed);
        for (int l = 0; l < fs.lanes; l++) cache_alloc_buffers(h, fs.slots, fs.group * cp->train_records, l);
        DCache &C = fs.C;
        C.depth = cp->cache_depth;
        // enough training candidates that about train_records of them reach the cache vertex
        C.train_stride = std::max(1, (int)((int64_t)fs.pixels * 2 / (3 * (int64_t)cp->train_records)));
        C.train_cap = cp->train_records;
        for (int a = 0; a < 3; a++) {
            C.lo[a] = h->bounds_lo[a];
            C.inv_ext[a] = 1.f / (h->bounds_hi[a] - h->bounds_lo[a]);
        }
        C.suffix = cp->suffix_len;
        // roughness a vertex's first lobe needs to end a path in the cache (Lambertian 1,
        // metal and glass 0, glossy sqrt(alpha^2)); TPT_CACHE_MIN_ROUGH in hundredths.
        // Textured room 960x540, 256 frames: thresholds 0.1 .. 0.5 lower relMSE from
        // 0.0161 to 0.0143 but take 33% longer per frame (paths go on through mirrors
        // and glass), so the default 0 ends every path at the cache depth
        C.min_rough = (float)env_int("TPT_CACHE_MIN_ROUGH", 0, 0, 100) / 100.f;
        // where the cache vertex's first lobe takes its colour from a texture, the network
        // learns radiance / albedo (TPT_NRC_ALBEDO: 0 nowhere, 1 textured lobes, 2 every
        // vertex).  Textured room 1920x1080, 4 runs each against a 16k-spp render: relMSE at
        // 256 frames 0.01796 / 0.01665 / 0.01732 (mean), at 64 frames 0.0435 / 0.0450 /
        // 0.0452; untextured scenes are unchanged by 1.  A moving average of the weights for
        // the queries (decay 0.95 / 0.99) lagged the still-improving network (-19% / -25%
        // bias at 16 frames) and did not lower the 256-frame error.
        C.albedo_factor = env_int("TPT_NRC_ALBEDO", 1, 0, 2);
        // probability of sampling the light mixture (instead of the BxDF) at the vertices
        // before the cache; the plain tracer keeps the reference's 0.2.  Textured room, three
        // runs each: 0.2 / 0.3 / 0.35 / 0.4 / 0.5 give 256-frame equal-error speedups of
        // 2.55 / 2.82 / 2.88 / 2.91 / 2.74 (the frames also end sooner: more paths reach
        // the emitters); city (sky, its only mixture light): 0.35 raised relMSE x time by 5%
        // and the bias from +3.6..4.0% to +4.3..6.2%, so mixtures holding the environment
        // keep 0.2.  TPT_CACHE_LIGHT_PROB (hundredths) sets it for every scene.
        const int lp_env = std::getenv("TPT_CACHE_LIGHT_PROB") ? env_int("TPT_CACHE_LIGHT_PROB", 20, 1, 99) : -1;
        C.light_prob = lp_env > 0 ? (float)lp_env / 100.f : (h->env_light ? 0.2f : 0.35f);
    }
    // With the cache ending every path at its cache vertex (depth cache_depth, any
    // material) and training paths suffix_len vertices later, no path is alive past
    // bounce cache_depth + suffix_len: the frame stops there instead of tracing
    // empty queues up to max_depth (same image).
    fs.last_bounce = fs.R.max_depth;
    if (fs.use_cache && fs.C.suffix > 0 && fs.C.min_rough == 0.f)
        fs.last_bounce = std::min(fs.R.max_depth, fs.C.depth + fs.C.suffix);
}

// The frame parameters with the staging buffers of one path pool.
DCache lane_cache(const DCache &C0, const CacheBufs &b) {
    DCache C = C0;
    C.feat = b.feat;
    C.feat_t = b.feat_t;
    C.rec_T = b.rec_T;
    C.rec_C = b.rec_C;
    C.q_slot = b.q_slot;
    C.t_slot = b.t_slot;
    C.counters = b.counters;
    return C;
}

// Enqueues frames [f0, f0 + n) in groups of up to fs.group frames, adding each
// frame to the per-pixel sums that the pools share (frame 0 starts them).
// Groups rotate over the lanes' pools and streams; a group's inference waits
// for the previous group's training and its accumulation for the previous
// accumulation, so the result equals one stream's.  Returns the number of groups.
// `ev`: 4 events per group bracketing the inference and the training launches,
// or null (room for n groups).  `clear_loss`: the last group's training loss is
// accumulated from zero (read back by the caller).
int enqueue_frames(tpt_hip_scene *h, FrameSetup &fs, PathBuffers *const pools[kMaxLanes], FrameChain &ch, int f0,
                   int n, hipEvent_t *ev, bool clear_loss) {
    const DScene &S = h->S;
    CacheState &c = h->cache;
    ch.create();
    if (fs.use_cache && !h->train_ev) HIP_OK(hipEventCreateWithFlags(&h->train_ev, hipEventDisableTiming));
    if (fs.lanes > 1) { // the other lanes start after the work already queued on the first
        HIP_OK(hipEventRecord(ch.start, h->stream));
        for (int l = 1; l < fs.lanes; l++) HIP_OK(hipStreamWaitEvent(lane_stream(h, l), ch.start, 0));
    }
    int groups = 0;
    for (int f = f0; f < f0 + n; groups++) {
        const int g = std::min(fs.group, f0 + n - f); // frames f .. f + g - 1
        const int lane = ch.next_lane;
        ch.next_lane = (lane + 1) % fs.lanes;
        ch.last_lane = lane;
        ch.last_frames = g;
        hipStream_t st = lane_stream(h, lane);
        DPaths &P = pools[lane]->P;
        CacheBufs &cb = c.buf[lane];
        hipEvent_t *e = ev ? &ev[4 * (size_t)groups] : nullptr;
        const int np = g * fs.pixels;
        const int grid = grid_for((size_t)np, kFrameBlock, h->cus);
        DCache C{};
        if (fs.use_cache) {
            C = lane_cache(fs.C, cb);
            C.frame = (uint32_t)(c.step + f);
            C.train_cap = g * fs.cp.train_records;
            launch_cache_reset(C, st);
        }
        launch_generate(S, fs.R, P, f, np, fs.base_mixed, kFrameBlock, st); // slot j: frame f + j / pixels
        int cur = 0;
        for (int b = 0; b <= fs.last_bounce; b++) {
            int32_t *q = cur == 0 ? P.queue_a : P.queue_b;
            int32_t *nx = cur == 0 ? P.queue_b : P.queue_a;
            launch_extend(S, P, q, cur, 1 - cur, grid, kFrameBlock, false, (uint64_t)np, st);
            launch_shade(S, fs.R, P, q, nx, cur, 1 - cur, grid, kFrameBlock, st, fs.use_cache ? &C : nullptr);
            launch_shadow(S, P, grid, kFrameBlock, false, (uint64_t)np, st);
            cur = 1 - cur;
        }
        if (fs.use_cache) {
            // inference reads the weights of the scene's latest training step, and
            // this group's training must follow it (any session or cached render)
            if (h->train_any) HIP_OK(hipStreamWaitEvent(st, h->train_ev, 0));
            if (clear_loss && f + g == f0 + n) HIP_OK(hipMemsetAsync(c.W.loss, 0, 4, st));
            if (e) HIP_OK(hipEventRecord(e[0], st));
            nrc_launch_infer(cb.feat, cb.q_slot, cb.counters + 0, np, c.W.wt, cb.pred, h->cus, st);
            if (e) HIP_OK(hipEventRecord(e[1], st));
            launch_cache_apply(P, C, cb.pred, np, st);
            launch_cache_targets(P, C, cb.target, st);
            if (e) HIP_OK(hipEventRecord(e[2], st));
            // train_steps Adam steps per frame, on consecutive mini-batches of
            // train_records / train_steps of the group's records
            const int K = fs.cp.train_steps, mb = (fs.cp.train_records + K - 1) / K;
            for (int k = 0; k < g * K; k++) {
                nrc_launch_train(cb.feat_t, cb.t_slot, cb.target, cb.counters + 1, C.train_cap, k * mb, mb, c.W, st);
                nrc_launch_adam(c.W, mb, fs.cp.learning_rate, ++c.step, st);
            }
            if (e) HIP_OK(hipEventRecord(e[3], st));
            HIP_OK(hipEventRecord(h->train_ev, st));
            h->train_any = true;
        }
        if (ch.any) HIP_OK(hipStreamWaitEvent(st, ch.acc, 0));
        launch_frame_accumulate(P, fs.pixels, fs.fp64, f == 0, g, st);
        HIP_OK(hipEventRecord(ch.acc, st));
        ch.any = true;
        f += g;
    }
    return groups;
}

} // namespace

int tpt_hip_render_cached(tpt_hip_scene *h, const tpt_render_params *p, const tpt_cache_params *cp, float *out_rgb,
                          tpt_cache_stats *stats) {
    if (!h || !p || !cp || !out_rgb) return set_err(TPT_ERR_INVALID_ARGUMENT, "null argument");
    if (cp->frames <= 0) return set_err(TPT_ERR_INVALID_ARGUMENT, "frames must be positive");
    try {
        auto t0 = std::chrono::steady_clock::now();
        FrameSetup fs;
        setup_frames(h, p, cp, fs);
        const int pixels = fs.pixels;
        ensure_paths(h, fs.slots, pixels);
        if (fs.lanes > 1) ensure_paths2(h, fs.slots, pixels);
        for (int l = 2; l < fs.lanes; l++) {
            PathBuffers &x = h->pbx[l - 2];
            if (!(x.mem && x.slots >= fs.slots && x.pixels == pixels)) {
                sync_lanes(h);
                alloc_paths(x, fs.slots, (size_t)pixels, false);
            }
            share_sums(x, h->pb);
        }
        PathBuffers *pools[kMaxLanes] = {&h->pb, &h->pb2, &h->pbx[0], &h->pbx[1]};
        hipStream_t st = h->stream;
        CacheState &c = h->cache;
        // four events per frame group (inference and training brackets), read after the last
        EventVec ev(fs.use_cache ? 4 * (size_t)cp->frames : 0);
        for (auto &x : ev) HIP_OK(hipEventCreate(&x));
        FrameChain ch;
        const int groups = enqueue_frames(h, fs, pools, ch, 0, cp->frames, ev.empty() ? nullptr : ev.data(), true);
        HIP_OK(hipStreamWaitEvent(st, ch.acc, 0)); // the last frame may have run on another lane's stream
        launch_frame_finalize(h->pb.P, pixels, fs.fp64, cp->frames, h->pb.out, st);
        HIP_OK(hipGetLastError());
        HIP_OK(hipMemcpyAsync(out_rgb, h->pb.out, (size_t)pixels * 12, hipMemcpyDeviceToHost, st));
        uint32_t cnt[2] = {0, 0};
        float loss = 0.f;
        if (fs.use_cache) {
            const CacheBufs &last = c.buf[ch.last_lane];
            HIP_OK(hipMemcpyAsync(cnt, last.counters, 8, hipMemcpyDeviceToHost, st));
            HIP_OK(hipMemcpyAsync(&loss, c.W.loss, 4, hipMemcpyDeviceToHost, st));
        }
        sync_lanes(h);
        double infer_ms = 0, train_ms = 0;
        for (size_t f = 0; !ev.empty() && f < (size_t)groups; f++) {
            float a = 0, b2 = 0;
            HIP_OK(hipEventElapsedTime(&a, ev[4 * f], ev[4 * f + 1]));
            HIP_OK(hipEventElapsedTime(&b2, ev[4 * f + 2], ev[4 * f + 3]));
            infer_ms += a;
            train_ms += b2;
        }
        if (stats) {
            std::memset(stats, 0, sizeof(*stats));
            stats->render_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
            stats->infer_ms = infer_ms;
            stats->train_ms = train_ms;
            // the last group's sums, per frame
            const int g = std::max(1, ch.last_frames);
            stats->last_loss = fs.use_cache ? loss / (g * cp->train_steps) : 0.0; // mean of the mini-batch means
            stats->queries = fs.use_cache ? cnt[0] / g : 0;
            stats->records = fs.use_cache ? std::min<uint32_t>(cnt[1], (uint32_t)(g * cp->train_records)) / g : 0;
            stats->steps = c.step;
        }
        return TPT_OK;
    } catch (const tpt_error &e) {
        return set_err(e.code, e.what());
    } catch (const std::exception &e) {
        return set_err(TPT_ERR_RUNTIME, e.what());
    }
}

// ---------------------------------------------------------------- progressive sessions
struct tpt_hip_progress {
    tpt_hip_scene *scene = nullptr;
    FrameSetup fs;
    PathBuffers pb;        // the session's own path pool and per-pixel sums
    PathBuffers lanes[3];  // the pools of frame lanes 1..3, adding into pb's sums
    FrameChain chain;
    int frames = 0;
    void sync() const {
        if (scene) sync_lanes(scene);
    }
    ~tpt_hip_progress() {
        if (scene) // nothing queued may still use the buffers
            for (int l = 0; l < kMaxLanes; l++) {
                hipStream_t x = lane_stream(scene, l);
                if (x) (void)hipStreamSynchronize(x);
            }
        if (pb.mem) (void)hipFree(pb.mem);
        for (PathBuffers &x : lanes)
            if (x.mem) (void)hipFree(x.mem);
    }
};

int tpt_hip_progress_create(tpt_hip_scene *scene, const tpt_render_params *params, const tpt_cache_params *cache,
                            tpt_hip_progress **out) {
    if (!scene || !params || !out) return set_err(TPT_ERR_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    tpt_hip_progress *s = nullptr;
    try {
        s = new tpt_hip_progress();
        setup_frames(scene, params, cache, s->fs);
        s->scene = scene;
        alloc_paths(s->pb, s->fs.slots, (size_t)s->fs.pixels, false);
        for (int l = 1; l < s->fs.lanes; l++) {
            alloc_paths(s->lanes[l - 1], s->fs.slots, (size_t)s->fs.pixels, false);
            share_sums(s->lanes[l - 1], s->pb);
        }
        *out = s;
        return TPT_OK;
    } catch (const tpt_error &e) {
        delete s;
        return set_err(e.code, e.what());
    } catch (const std::exception &e) {
        delete s;
        return set_err(TPT_ERR_RUNTIME, e.what());
    }
}

int tpt_hip_progress_add(tpt_hip_progress *s, int32_t frames) {
    if (!s) return set_err(TPT_ERR_INVALID_ARGUMENT, "null argument");
    if (frames <= 0) return set_err(TPT_ERR_INVALID_ARGUMENT, "frames must be positive");
    if ((int64_t)s->frames + frames > INT32_MAX) return set_err(TPT_ERR_INVALID_ARGUMENT, "too many frames");
    try {
        HIP_OK(hipSetDevice(s->scene->device));
        PathBuffers *pools[kMaxLanes] = {&s->pb, &s->lanes[0], &s->lanes[1], &s->lanes[2]};
        enqueue_frames(s->scene, s->fs, pools, s->chain, s->frames, frames, nullptr, false);
        HIP_OK(hipGetLastError());
        s->frames += frames;
        return TPT_OK;
    } catch (const tpt_error &e) {
        return set_err(e.code, e.what());
    } catch (const std::exception &e) {
        return set_err(TPT_ERR_RUNTIME, e.what());
    }
}

int tpt_hip_progress_wait(tpt_hip_progress *s) {
    if (!s) return set_err(TPT_ERR_INVALID_ARGUMENT, "null argument");
    try {
        HIP_OK(hipSetDevice(s->scene->device));
        s->sync();
        return TPT_OK;
    } catch (const tpt_error &e) {
        return set_err(e.code, e.what());
    }
}

int tpt_hip_progress_read(tpt_hip_progress *s, float *out_rgb, int32_t *frames) {
    if (!s || !out_rgb) return set_err(TPT_ERR_INVALID_ARGUMENT, "null argument");
    if (s->frames == 0) return set_err(TPT_ERR_INVALID_ARGUMENT, "no frames rendered yet");
    try {
        HIP_OK(hipSetDevice(s->scene->device));
        hipStream_t st = s->scene->stream;
        HIP_OK(hipStreamWaitEvent(st, s->chain.acc, 0)); // the last frame may have run on another lane's stream
        launch_frame_finalize(s->pb.P, s->fs.pixels, s->fs.fp64, s->frames, s->pb.out, st);
        HIP_OK(hipGetLastError());
        HIP_OK(hipMemcpyAsync(out_rgb, s->pb.out, (size_t)s->fs.pixels * 12, hipMemcpyDeviceToHost, st));
        HIP_OK(hipStreamSynchronize(st));
        if (frames) *frames = s->frames;
        return TPT_OK;
    } catch (const tpt_error &e) {
        return set_err(e.code, e.what());
    } catch (const std::exception &e) {
        return set_err(TPT_ERR_RUNTIME, e.what());
    }
}

void tpt_hip_progress_destroy(tpt_hip_progress *s) { delete s; }

int tpt_hip_cache_weights(tpt_hip_scene *h, const float *set, float *get) {
    if (!h) return set_err(TPT_ERR_INVALID_ARGUMENT, "null scene");
    try {
        HIP_OK(hipSetDevice(h->device));
        cache_alloc_weights(h);
        if (set) cache_set_weights(h, set);
        else if (h->cache.step == 0 && !get) cache_init_weights(h, 0);
        if (get) {
            sync_lanes(h); // frames on any lane may still train the network
            HIP_OK(hipMemcpy(get, h->cache.W.w, 4 * kNrcParams, hipMemcpyDeviceToHost));
        }
        return TPT_OK;
    } catch (const tpt_error &e) {
        return set_err(e.code, e.what());
    }
}

namespace {
// Host records -> device staging (rows 0..n-1 of the feature buffer, targets).
void cache_stage(tpt_hip_scene *h, const uint16_t *inputs, const float *targets, int n) {
    cache_alloc_weights(h);
    cache_alloc_buffers(h, (size_t)std::max(n, 64), std::max(n, 64), 0);
    CacheBufs &b = h->cache.buf[0];
    sync_lanes(h); // frames on any lane may still use lane 0's buffers or train the network
    HIP_OK(hipMemcpy(b.feat, inputs, (size_t)n * kNrcInputs * 2, hipMemcpyHostToDevice));
    if (targets) {
        std::vector<float4> t(n);
        for (int i = 0; i < n; i++) t[i] = make_float4(targets[3 * i], targets[3 * i + 1], targets[3 * i + 2], 0.f);
        HIP_OK(hipMemcpy(b.target, t.data(), (size_t)n * 16, hipMemcpyHostToDevice));
    }
    uint32_t cnt[2] = {(uint32_t)n, (uint32_t)n};
    HIP_OK(hipMemcpy(b.counters, cnt, 8, hipMemcpyHostToDevice));
}
} // namespace

int tpt_hip_cache_eval(tpt_hip_scene *h, const uint16_t *inputs, int32_t n, float *out) {
    if (!h || !inputs || !out || n <= 0) return set_err(TPT_ERR_INVALID_ARGUMENT, "null argument");
    try {
        HIP_OK(hipSetDevice(h->device));
        cache_stage(h, inputs, nullptr, n);
        CacheState &c = h->cache;
        const CacheBufs &b = c.buf[0];
        nrc_launch_infer(b.feat, nullptr, b.counters, n, c.W.wt, b.pred, h->cus, h->stream);
        HIP_OK(hipGetLastError());
        std::vector<float4> o(n);
        HIP_OK(hipMemcpyAsync(o.data(), b.pred, (size_t)n * 16, hipMemcpyDeviceToHost, h->stream));
        HIP_OK(hipStreamSynchronize(h->stream));
        for (int i = 0; i < n; i++) {
            out[3 * i] = o[i].x;
            out[3 * i + 1] = o[i].y;
            out[3 * i + 2] = o[i].z;
        }
        return TPT_OK;
    } catch (const tpt_error &e) {
        return set_err(e.code, e.what());
    }
}

int tpt_hip_cache_train(tpt_hip_scene *h, const uint16_t *inputs, const float *targets, int32_t n, float lr,
                        float *loss, float *grad) {
    if (!h || !inputs || !targets || n <= 0) return set_err(TPT_ERR_INVALID_ARGUMENT, "null argument");
    try {
        HIP_OK(hipSetDevice(h->device));
        cache_stage(h, inputs, targets, n);
        CacheState &c = h->cache;
        const CacheBufs &b = c.buf[0];
        hipStream_t st = h->stream;
        HIP_OK(hipMemsetAsync(c.W.loss, 0, 4, st));
        nrc_launch_train(b.feat, nullptr, b.target, b.counters + 1, n, 0, n, c.W, st);
        nrc_launch_adam(c.W, n, lr, ++c.step, st); // sums the step's gradient into grad
        HIP_OK(hipGetLastError());
        if (grad) HIP_OK(hipMemcpyAsync(grad, c.W.grad, 4 * kNrcParams, hipMemcpyDeviceToHost, st));
        if (loss) HIP_OK(hipMemcpyAsync(loss, c.W.loss, 4, hipMemcpyDeviceToHost, st));
        HIP_OK(hipStreamSynchronize(st));
        return TPT_OK;
    } catch (const tpt_error &e) {
        return set_err(e.code, e.what());
    }
}

static int trace_common(tpt_hip_scene *h, const float *rays, uint64_t n, void *out_a, void *out_b, uint32_t flags,
                        void *stream, bool closest) {
    if (!h || (!rays && n)) return set_err(TPT_ERR_INVALID_ARGUMENT, "null argument");
    if (n == 0) return TPT_OK;
    try {
        HIP_OK(hipSetDevice(h->device));
        hipStream_t st = (flags & TPT_PTR_DEVICE) ? (hipStream_t)stream : h->stream;
        const float4 *d_rays;
        void *d_a, *d_b = nullptr;
        void *tmp = nullptr;
        bool dev = (flags & TPT_PTR_DEVICE) != 0;
        bool count = (flags & TPT_COUNT_TRAFFIC) != 0;
        if (dev) {
            d_rays = (const float4 *)rays;
            d_a = out_a;
            d_b = out_b;
        } else {
            size_t bytes_r = n * 32, bytes_a = closest ? n * 4 : n, bytes_b = closest ? n * 4 : 0;
            HIP_OK(hipMalloc(&tmp, align256(bytes_r) + align256(bytes_a) + align256(bytes_b) + 256));
            d_rays = (const float4 *)tmp;
            d_a = (uint8_t *)tmp + align256(bytes_r);
            d_b = (uint8_t *)d_a + align256(bytes_a);
            HIP_OK(hipMemcpyAsync((void *)d_rays, rays, bytes_r, hipMemcpyHostToDevice, st));
        }
        HIP_OK(hipMemsetAsync(h->trace_counters, 0, 64, st)); // traffic + persistent fetch counters
        int block = 256;
        int grid = grid_for(n, block, h->cus);
        if (closest)
            launch_trace_closest(h->S, d_rays, n, (int32_t *)d_a, (float *)d_b, h->trace_counters, count, grid, block,
                                 st);
        else
            launch_trace_occluded(h->S, d_rays, n, (uint8_t *)d_a, h->trace_counters, grid, block, st);
        HIP_OK(hipGetLastError());
        if (!dev) {
            if (closest) {
                HIP_OK(hipMemcpyAsync(out_a, d_a, n * 4, hipMemcpyDeviceToHost, st));
                HIP_OK(hipMemcpyAsync(out_b, d_b, n * 4, hipMemcpyDeviceToHost, st));
            } else {
                HIP_OK(hipMemcpyAsync(out_a, d_a, n, hipMemcpyDeviceToHost, st));
            }
            HIP_OK(hipStreamSynchronize(st));
            HIP_OK(hipFree(tmp));
        }
        if (count) {
            uint64_t c[4];
            HIP_OK(hipMemcpyAsync(c, h->trace_counters, 32, hipMemcpyDeviceToHost, st));
            HIP_OK(hipStreamSynchronize(st));
            h->last_nodes = c[0];
            h->last_tris = c[1];
        }
        return TPT_OK;
    } catch (const tpt_error &e) {
        return set_err(e.code, e.what());
    } catch (const std::exception &e) {
        return set_err(TPT_ERR_RUNTIME, e.what());
    }
}

int tpt_hip_trace_closest(tpt_hip_scene *h, const float *rays, uint64_t n, int32_t *prim_id, float *t, uint32_t flags,
                          void *stream) {
    return trace_common(h, rays, n, prim_id, t, flags, stream, true);
}

int tpt_hip_trace_occluded(tpt_hip_scene *h, const float *rays, uint64_t n, uint8_t *occluded, uint32_t flags,
                           void *stream) {
    return trace_common(h, rays, n, occluded, nullptr, flags, stream, false);
}

int tpt_hip_last_trace_counters(tpt_hip_scene *h, uint64_t *nodes, uint64_t *tris) {
    if (!h) return set_err(TPT_ERR_INVALID_ARGUMENT, "null scene");
    if (nodes) *nodes = h->last_nodes;
    if (tris) *tris = h->last_tris;
    return TPT_OK;
}

} // extern "C"
