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
#include <string>
#include <vector>

#include "../../include/tinypt_hip.h"
#include "dev_math.h"
#include "dev_types.h"
#include "scene_build.h"

namespace tpt {
void launch_generate(const DScene &, const DRender &, const DPaths &, int, int, uint64_t, int, hipStream_t);
void launch_extend(const DScene &, const DPaths &, const int32_t *, int, int, int, int, bool, hipStream_t);
void launch_shade(const DScene &, const DRender &, const DPaths &, const int32_t *, int32_t *, int, int, int, int,
                  hipStream_t);
void launch_shadow(const DScene &, const DPaths &, int, int, bool, hipStream_t);
void launch_accumulate(const DRender &, const DPaths &, int, int, int, hipStream_t);
void launch_finalize(const DRender &, const DPaths &, float *, hipStream_t);
void launch_trace_closest(const DScene &, const float4 *, uint64_t, int32_t *, float *, uint32_t *, bool, int, int,
                          hipStream_t);
void launch_trace_occluded(const DScene &, const float4 *, uint64_t, uint8_t *, uint32_t *, int, int, hipStream_t);
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

} // namespace

struct tpt_hip_scene {
    int device = 0;
    int cus = 256; // compute units of the device (grid sizing)
    void *arena = nullptr;
    size_t arena_bytes = 0;
    DScene S{};
    uint64_t num_nodes = 0, num_tris = 0, num_analytic = 0;
    PathBuffers pb;
    uint32_t *trace_counters = nullptr; // 8 x u32
    uint64_t last_nodes = 0, last_tris = 0;
    hipStream_t stream = nullptr;
    ~tpt_hip_scene() {
        if (pb.mem) (void)hipFree(pb.mem);
        if (arena) (void)hipFree(arena);
        if (trace_counters) (void)hipFree(trace_counters);
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
    S.layout = H.layout;
    const char *pers = std::getenv("TPT_PERSISTENT");
    // traversal kernels: 0 = one ray per thread, 1 = persistent with per-wave refill,
    // 2 (default) = persistent + parked leaves
    S.persistent = 2;
    if (pers && std::string(pers) == "0") S.persistent = 0;
    if (pers && std::string(pers) == "1") S.persistent = 1;
    // traversal stack: private scratch by default; TPT_LDS_STACK=1 stages it in LDS
    const char *lds = std::getenv("TPT_LDS_STACK");
    S.lds_stack = (lds && std::string(lds) == "1") ? 1 : 0;
    // persistent traversal workgroups and the LDS node cache of the quantized layout
    // (the top of the tree, in BFS order at the start of the node array)
    S.trav_block = env_int("TPT_TRAV_BLOCK", kTravBlock, 64, 1024) / 64 * 64;
    S.lds_nodes = H.layout == kNodeBvh4Q ? std::min(H.bfs_prefix, env_int("TPT_LDS_NODES", kLdsNodes, 0, 2048)) : 0;
    S.num_nodes = (int32_t)(H.nodes.size() / kNodeFloat4[H.layout]);
    S.root = H.root;
    S.num_tris = (int32_t)(H.tris_bvh.size() / 3);
    S.num_analytic = (int32_t)H.analytic.size();
    S.num_delta = (int32_t)H.delta.size();
    S.num_lights = (int32_t)H.lights.size();
    S.env_w = H.env_w;
    S.env_h = H.env_h;
    S.bg_map = H.bg_map;
    S.bg = H.bg;
    std::memcpy(S.bg_rot, H.bg_rot, sizeof(S.bg_rot));
    S.cam = H.cam;
    h->num_nodes = H.root >= 0 && !(H.root & (int32_t)kLeafBit) ? (uint64_t)S.num_nodes : 0;
    h->num_tris = H.root == -1 ? 0 : (uint64_t)S.num_tris;
    h->num_analytic = H.analytic.size();
}

void ensure_paths(tpt_hip_scene *h, size_t slots, int pixels) {
    PathBuffers &pb = h->pb;
    if (pb.mem && pb.slots >= slots && pb.pixels == pixels) return;
    if (pb.mem) {
        HIP_OK(hipFree(pb.mem));
        pb.mem = nullptr;
    }
    size_t off = 0;
    auto take = [&](size_t bytes) {
        size_t o = off;
        off = align256(off + bytes);
        return o;
    };
    size_t o_org = take(slots * 16), o_dir = take(slots * 16), o_dec = take(slots * 16), o_col = take(slots * 16);
    size_t o_rng = take(slots * 8), o_key = take(slots * 8), o_dep = take(slots * 4);
    size_t o_qa = take(slots * 4), o_qb = take(slots * 4), o_ss = take(slots * 4);
    size_t o_so = take(slots * 16), o_sd = take(slots * 16), o_sv = take(slots * 16), o_hit = take(slots * 16);
    size_t o_cnt = take(64);
    size_t o_sf = take((size_t)pixels * 12), o_sdd = take((size_t)pixels * 24), o_pf = take((size_t)pixels * 12),
           o_pd = take((size_t)pixels * 24), o_out = take((size_t)pixels * 12);
    HIP_OK(hipMalloc(&pb.mem, off));
    pb.bytes = off;
    pb.slots = slots;
    pb.pixels = pixels;
    uint8_t *b = (uint8_t *)pb.mem;
    DPaths &P = pb.P;
    P.org = (float4 *)(b + o_org);
    P.dir = (float4 *)(b + o_dir);
    P.decay = (float4 *)(b + o_dec);
    P.color = (float4 *)(b + o_col);
    P.rng = (uint64_t *)(b + o_rng);
    P.key = (uint64_t *)(b + o_key);
    P.depth = (int32_t *)(b + o_dep);
    P.queue_a = (int32_t *)(b + o_qa);
    P.queue_b = (int32_t *)(b + o_qb);
    P.shadow_slot = (int32_t *)(b + o_ss);
    P.shadow_org = (float4 *)(b + o_so);
    P.shadow_dir = (float4 *)(b + o_sd);
    P.shadow_val = (float4 *)(b + o_sv);
    P.hit = (float4 *)(b + o_hit);
    P.counters = (uint32_t *)(b + o_cnt);
    P.sub_f = (float *)(b + o_sf);
    P.sub_d = (double *)(b + o_sdd);
    P.pix_f = (float *)(b + o_pf);
    P.pix_d = (double *)(b + o_pd);
    pb.out = (float *)(b + o_out);
    HIP_OK(hipMemset(P.counters, 0, 64));
}

int block_size_for(int spp) { // PixelSampler ctor (cpu/path_tracer.cpp:10-21)
    const int m = 16;
    if (spp < 2 * 2 * m) return 1;
    if (spp < 3 * 3 * m) return 2;
    if (spp < 4 * 4 * m) return 3;
    return 4;
}

int grid_for(size_t n, int block, int cus = 256) {
    size_t g = (n + block - 1) / block;
    // grid-stride / persistent kernels: 8 resident blocks of 256 threads (32 waves) per CU
    size_t cap = (size_t)cus * 2048 / block;
    if (g > cap) g = cap;
    if (g < 1) g = 1;
    return (int)g;
}

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
        int stack_bytes = h->S.lds_stack ? kLdsStack * h->S.trav_block * 4 : 0;
        if (stack_bytes > max_shared) {
            h->S.lds_stack = 0;
            stack_bytes = 0;
        }
        h->S.lds_nodes = std::min(h->S.lds_nodes, (max_shared - stack_bytes) / 64);
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
        DRender R{};
        R.block = block_size_for(p->spp);
        R.nsub = p->spp / (R.block * R.block);
        R.spp_eff = R.nsub * R.block * R.block;
        R.seed = p->seed;
        R.max_depth = p->max_depth;
        R.rr_depth = p->rr_depth;
        R.shadow_bounded = p->shadow_bounded;
        R.output = p->output == TPT_OUT_HDR_MEAN ? 1 : 0;
        R.accum_fp64 = p->accum_fp64 ? 1 : 0;
        R.pixels = pixels;
        size_t cap = p->queue_capacity > 0 ? (size_t)p->queue_capacity : ((size_t)1 << 24);
        int k = (int)std::max<size_t>(1, std::min<size_t>((size_t)R.spp_eff, cap / (size_t)pixels));
        R.k = k;
        int block = p->block_size > 0 ? p->block_size : 256;
        if (block % 64 != 0 || block > 1024) return set_err(TPT_ERR_INVALID_ARGUMENT, "block_size must be a multiple of 64 <= 1024");
        ensure_paths(h, (size_t)pixels * k, pixels);
        DPaths &P = h->pb.P;
        hipStream_t st = h->stream;
        uint64_t base_mixed = splitmix64(p->seed);
        bool timing = stats != nullptr;
        bool count = p->count_traffic != 0;
        if (count) HIP_OK(hipMemsetAsync(P.counters + 4, 0, 48, st));
        std::vector<hipEvent_t> evs;
        auto ev = [&]() {
            hipEvent_t e;
            HIP_OK(hipEventCreate(&e));
            HIP_OK(hipEventRecord(e, st));
            evs.push_back(e);
            return (int)evs.size() - 1;
        };
        std::vector<std::pair<int, int>> ext_ev, shade_ev, shadow_ev;
        int first = 1;
        for (int wave_base = 0; wave_base < R.spp_eff; wave_base += k) {
            int k_active = std::min(k, R.spp_eff - wave_base);
            int n = pixels * k_active;
            launch_generate(S, R, P, wave_base, n, base_mixed, block, st);
            int grid = grid_for((size_t)n, block, h->cus);
            int cur = 0;
            for (int b = 0; b <= R.max_depth; b++) {
                int32_t *q = cur == 0 ? P.queue_a : P.queue_b;
                int32_t *nx = cur == 0 ? P.queue_b : P.queue_a;
                int e0 = timing ? ev() : -1;
                launch_extend(S, P, q, cur, 1 - cur, grid, block, count, st);
                int e1 = timing ? ev() : -1;
                launch_shade(S, R, P, q, nx, cur, 1 - cur, grid, block, st);
                int e2 = timing ? ev() : -1;
                launch_shadow(S, P, grid, block, count, st);
                int e3 = timing ? ev() : -1;
                if (timing) {
                    ext_ev.push_back({e0, e1});
                    shade_ev.push_back({e1, e2});
                    shadow_ev.push_back({e2, e3});
                }
                cur = 1 - cur;
            }
            launch_accumulate(R, P, wave_base, k_active, first, st);
            first = 0;
        }
        launch_finalize(R, P, h->pb.out, st);
        HIP_OK(hipGetLastError());
        HIP_OK(hipMemcpyAsync(out_rgb, h->pb.out, (size_t)pixels * 12, hipMemcpyDeviceToHost, st));
        HIP_OK(hipStreamSynchronize(st));
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
                uint64_t c[4];
                HIP_OK(hipMemcpy(c, P.counters + 4, 32, hipMemcpyDeviceToHost));
                stats->nodes_visited = c[0];
                stats->tris_tested = c[1];
                stats->rays_closest = c[2];
                stats->rays_shadow = c[3];
            }
        }
        for (auto e : evs) (void)hipEventDestroy(e);
        return TPT_OK;
    } catch (const tpt_error &e) {
        return set_err(e.code, e.what());
    } catch (const std::exception &e) {
        return set_err(TPT_ERR_RUNTIME, e.what());
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
