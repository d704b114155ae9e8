// kernels.hip — gfx950 wavefront path tracer (device=hip).
//
// Stages per wave of paths (all state SoA in HBM, queues compacted with wave64
// ballot + one atomic per wavefront):
//   generate   pixel/sample -> camera ray + erand48 state        (cpu/path_tracer.cpp:77-96, camera.cpp:23-34)
//   extend     closest hit over analytic prims + the triangle BVH (cpu/scene.h:27, object.cpp:389-411, bvh.cpp:78-118)
//   shade      emission, BxDF/light mixture sampling, delta-light
//              shadow-ray emission, throughput update, RR         (cpu/path_tracer.cpp:102-168)
//   shadow     any-hit for delta-light shadow rays                (cpu/path_tracer.cpp:129-139)
//   accumulate per-pixel sub-pixel sums in sample order           (cpu/path_tracer.cpp:80-96)
//   finalize   clip(sum / b^2) into the output image
#include <hip/hip_runtime.h>

#include <type_traits>

#include "dev_scene.h"

namespace tpt {

// ---------------------------------------------------------------- traversal
// Child box slab test.  The GPU BVH is this library's own tree, so the test
// only has to be conservative: a zero direction component uses a large finite
// reciprocal (no 0*inf NaNs) and the far distance is widened by 2^-17
// relative so that a triangle whose exact Moller-Trumbore t lies on a box face
// is never culled.
__device__ __forceinline__ float safe_rcp(float d) {
    return (fabsf(d) > 1e-30f) ? 1.f / d : copysignf(1e30f, d);
}
// Per-ray constants of the slab test: the origin and the reciprocal direction.
// Slab distances are (l - o) * inv, exactly as the reference's AABB::intersect
// (cpu/bvh.cpp:8-27) computes them; the far distance is widened by 2^-17
// relative, the near distance is clamped to EPS and the culling bound is the
// current closest t, all folded into max3 / min3.  (A fma(l, inv, -o*inv)
// form saves 6 ops per box but needs two more live registers per lane and
// measured 25% slower on gfx950 through spills.)
struct RayPre {
    V3 o, inv;
};
__device__ __forceinline__ RayPre ray_pre(V3 o, V3 d) {
    RayPre r;
    r.o = o;
    r.inv = V3{safe_rcp(d.x), safe_rcp(d.y), safe_rcp(d.z)};
    return r;
}
static constexpr float kWiden = 1.f + 0x1p-17f;
__device__ __forceinline__ float cull_bound(const RayPre &, float tmax) { return tmax; }
__device__ __forceinline__ bool box_hit(float lx, float ly, float lz, float hx, float hy, float hz, const RayPre &r,
                                        float tw, float &tnear) {
    float t0x = (lx - r.o.x) * r.inv.x, t1x = (hx - r.o.x) * r.inv.x;
    float t0y = (ly - r.o.y) * r.inv.y, t1y = (hy - r.o.y) * r.inv.y;
    float t0z = (lz - r.o.z) * r.inv.z, t1z = (hz - r.o.z) * r.inv.z;
    float n = fmaxf(fmaxf(fminf(t0x, t1x), fminf(t0y, t1y)), fmaxf(fminf(t0z, t1z), kEPS * 0.99f));
    float f = fminf(fminf(fmaxf(t0x, t1x), fmaxf(t0y, t1y)), fmaxf(t0z, t1z));
    tnear = n;
    return n <= fminf(f * kWiden, tw);
}

struct AlphaCtx {
    uint64_t key;
    int depth;
    bool on;
};

struct HitRec {
    float t;
    int32_t gid;
    float u, v; // triangle barycentrics; analytic: u = into
};

// alpha test of a triangle candidate (object.cpp:273-287)
__device__ __forceinline__ bool tri_alpha_pass(const DScene &S, int tri, float u, float v, const AlphaCtx &ac,
                                               int32_t gid) {
    int4 i0 = S.tri_info[2 * tri], i1 = S.tri_info[2 * tri + 1];
    const DMaterial &m = S.materials[i0.x];
    V2 uv;
    if (i1.x >= 0) {
        const float *tc = S.texcoords;
        float w0 = 1.f - u - v;
        uv = {(w0 * tc[2 * i1.x] + u * tc[2 * i1.y]) + v * tc[2 * i1.z],
              (w0 * tc[2 * i1.x + 1] + u * tc[2 * i1.y + 1]) + v * tc[2 * i1.z + 1]};
    } else {
        uv = {u, v};
    }
    return !(alpha_at(S, m.alpha, uv) < alpha_draw(ac.key, ac.depth, gid));
}

// Dynamic LDS of the persistent traversal kernels: [cached top nodes: 64 B each]
// [stack: kLdsStack * blockDim.x entries when the LDS stack is on].
extern __shared__ int32_t tpt_lds_stack[];

// Traversal stack: with LDS on, the first kLdsStack entries of every lane live
// in LDS, entry-major ([entry][lane]), so a lane always touches its own bank
// whatever its depth; deeper entries (rare: the host bounds the worst case by
// kStackSize) spill to a per-lane private array.  With LDS off the whole stack
// is the per-lane private array (scratch accesses are lane-interleaved by the
// hardware, so they coalesce).
// LDS data is addressed through address-space-3 pointers: with a generic
// pointer the compiler merges the LDS and the private store of set() into one
// flat store through a selected address, which pays both the vector-memory and
// the LDS path.
typedef float v4f __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(3))) int32_t lds_i32;
typedef __attribute__((address_space(3))) v4f lds_f4;
__device__ __forceinline__ float4 from_v4f(v4f v) { return make_float4(v.x, v.y, v.z, v.w); }
template <bool LDS>
struct TStack {
    lds_i32 *lds;
    int stride, lane;
    int32_t spill[kStackSize];
    // `cached_nodes`: 64-B node records stored in front of the stack
    __device__ __forceinline__ void init(int cached_nodes) {
        lds = (lds_i32 *)tpt_lds_stack + 16 * cached_nodes;
        stride = blockDim.x;
        lane = threadIdx.x;
    }
    __device__ __forceinline__ void set(int sp, int32_t v) {
        if (LDS && sp < kLdsStack) lds[sp * stride + lane] = v;
        else spill[sp] = v;
    }
    __device__ __forceinline__ int32_t get(int sp) const {
        if (LDS && sp < kLdsStack) return lds[sp * stride + lane];
        return spill[sp];
    }
};

// Leaf: Moller-Trumbore on each triangle record (48 B) of the leaf range.
__device__ __forceinline__ void leaf_tests(const DScene &S, int32_t cur, V3 o, V3 d, const AlphaCtx &ac, HitRec &best,
                                           bool any, uint32_t &tris, bool count, bool &done) {
    uint32_t ref = (uint32_t)cur;
    int first = (int)((ref & ~kLeafBit) >> 3);
    int cnt = (int)(ref & 7u) + 1;
    for (int k = 0; k < cnt; k++) {
        const float4 *tr = S.tris_bvh + 3 * (first + k);
        float4 r0 = tr[0], r1 = tr[1], r2 = tr[2];
        if (count) tris++;
        float t, u, v;
        if (!tri_test(mk4(r0), mk4(r1), mk4(r2), o, d, t, u, v)) continue;
        int32_t gid = __float_as_int(r0.w);
        if (kEPS <= t && (t < best.t || (t == best.t && gid < best.gid))) {
            if (!any && ac.on && (__float_as_int(r2.w) & 1)) {
                if (!tri_alpha_pass(S, __float_as_int(r1.w), u, v, ac, gid)) continue;
            }
            best = HitRec{t, gid, u, v};
            if (any) {
                done = true;
                return;
            }
        }
    }
}

// Analytic primitives (the reference's linear top level, cpu/object.cpp:398-405).
// Returns true when an any-hit query is already answered.
template <bool ANY>
__device__ __forceinline__ bool analytic_tests(const DScene &S, V3 o, V3 d, const AlphaCtx &ac, HitRec &best) {
    for (int i = 0; i < S.num_analytic; i++) {
        const DAnalytic &p = S.analytic[i];
        float t;
        bool into = true;
        if (p.type == A_SPHERE) {
            t = sphere_t(p, o, d, into);
        } else {
            V2 uv;
            t = plane_t(p, o, d, kINF, uv);
        }
        if (t < kINF && (t < best.t || (t == best.t && p.gid < best.gid))) {
            if (!ANY && ac.on && p.type == A_SPHERE && !S.materials[p.material].alpha_opaque) {
                V3 n = nrm(sub(at(o, d, t), mk(p.center)));
                V2 uv{(atan2f(n.y, n.x) + kPI) * (.5f * kINV_PI), acosf(n.z) * kINV_PI};
                if (alpha_at(S, S.materials[p.material].alpha, uv) < alpha_draw(ac.key, ac.depth, p.gid)) continue;
            }
            best = HitRec{t, p.gid, into ? 1.f : 0.f, 0.f};
            if (ANY) return true;
        }
    }
    return false;
}

// Slab test on precomputed slab distances (t0 / t1 per axis, any order).
__device__ __forceinline__ bool slab_hit(float t0x, float t1x, float t0y, float t1y, float t0z, float t1z, float tw,
                                         float &tnear) {
    float n = fmaxf(fmaxf(fminf(t0x, t1x), fminf(t0y, t1y)), fmaxf(fminf(t0z, t1z), kEPS * 0.99f));
    float f = fminf(fminf(fmaxf(t0x, t1x), fmaxf(t0y, t1y)), fmaxf(t0z, t1z));
    tnear = n;
    return n <= fminf(f * kWiden, tw);
}

// Inner-node visit: tests the node's child boxes against the ray (culling at the
// current closest t `tw`), pushes the hit children except the nearest on the
// stack (farthest first) and returns the nearest in `nearc`; false when no child
// is hit.  NL selects the node layout (dev_types.h).
// `cache`: quantized nodes [0, cache) are read from the workgroup's LDS copy
// (persistent traversal kernels), the others from the node array.
template <int NL, class Stk>
__device__ __forceinline__ bool visit_inner(const DScene &S, int32_t cur, const RayPre &rp, float tw, int32_t &nearc,
                                            int &sp, Stk &stk, int cache) {
    if (NL == kNodeBvh2) {
        const float4 *nd = S.nodes + 4 * cur;
        float4 a = nd[0], b = nd[1], c = nd[2], e = nd[3];
        float n0, n1;
        bool h0 = box_hit(a.x, a.y, a.z, a.w, b.x, b.y, rp, tw, n0);
        bool h1 = box_hit(b.z, b.w, c.x, c.y, c.z, c.w, rp, tw, n1);
        int32_t c0 = __float_as_int(e.x), c1 = __float_as_int(e.y);
        if (h0 && h1) {
            bool swap = n1 < n0;
            stk.set(sp++, swap ? c0 : c1);
            nearc = swap ? c1 : c0;
            return true;
        }
        nearc = h0 ? c0 : c1;
        return h0 || h1;
    }
    float tn[4];
    int32_t ch[4];
    bool hb[4];
    if (NL == kNodeBvh4) {
        // 128-B node: f32 child boxes, 7 dwordx4 loads issued together (branch-free tests)
        const float4 *nd = S.nodes + 8 * cur;
        float4 lx = nd[0], hx = nd[1], ly = nd[2], hy = nd[3], lz = nd[4], hz = nd[5], rf = nd[6];
        ch[0] = __float_as_int(rf.x);
        ch[1] = __float_as_int(rf.y);
        ch[2] = __float_as_int(rf.z);
        ch[3] = __float_as_int(rf.w);
        hb[0] = box_hit(lx.x, ly.x, lz.x, hx.x, hy.x, hz.x, rp, tw, tn[0]);
        hb[1] = box_hit(lx.y, ly.y, lz.y, hx.y, hy.y, hz.y, rp, tw, tn[1]);
        hb[2] = box_hit(lx.z, ly.z, lz.z, hx.z, hy.z, hz.z, rp, tw, tn[2]);
        hb[3] = box_hit(lx.w, ly.w, lz.w, hx.w, hy.w, hz.w, rp, tw, tn[3]);
    } else {
        // 64-B quantized node (4 dwordx4 loads): bound = origin + q * scale per axis, so
        // a slab distance is fma(q, scale * inv, (origin - o) * inv)
        float4 n0, n1, n2, n3;
        if (cur < cache) {
            const lds_f4 *nd = (const lds_f4 *)tpt_lds_stack + 4 * cur;
            n0 = from_v4f(nd[0]);
            n1 = from_v4f(nd[1]);
            n2 = from_v4f(nd[2]);
            n3 = from_v4f(nd[3]);
        } else {
            const float4 *nd = S.nodes + 4 * cur;
            n0 = nd[0];
            n1 = nd[1];
            n2 = nd[2];
            n3 = nd[3];
        }
        float ax = n0.w * rp.inv.x, ay = n1.x * rp.inv.y, az = n1.y * rp.inv.z;
        float bx = (n0.x - rp.o.x) * rp.inv.x, by = (n0.y - rp.o.y) * rp.inv.y, bz = (n0.z - rp.o.z) * rp.inv.z;
        uint32_t qlx = __float_as_uint(n1.z), qhx = __float_as_uint(n1.w);
        uint32_t qly = __float_as_uint(n2.x), qhy = __float_as_uint(n2.y);
        uint32_t qlz = __float_as_uint(n2.z), qhz = __float_as_uint(n2.w);
        ch[0] = __float_as_int(n3.x);
        ch[1] = __float_as_int(n3.y);
        ch[2] = __float_as_int(n3.z);
        ch[3] = __float_as_int(n3.w);
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int sh = 8 * i;
            hb[i] = slab_hit(fmaf((float)((qlx >> sh) & 0xffu), ax, bx), fmaf((float)((qhx >> sh) & 0xffu), ax, bx),
                             fmaf((float)((qly >> sh) & 0xffu), ay, by), fmaf((float)((qhy >> sh) & 0xffu), ay, by),
                             fmaf((float)((qlz >> sh) & 0xffu), az, bz), fmaf((float)((qhz >> sh) & 0xffu), az, bz), tw,
                             tn[i]);
        }
    }
#pragma unroll
    for (int i = 0; i < 4; i++) {
        hb[i] = hb[i] & (ch[i] != kEmptyChild);
        tn[i] = hb[i] ? tn[i] : kINF;
    }
    int nh = (int)hb[0] + (int)hb[1] + (int)hb[2] + (int)hb[3];
    if (nh == 0) return false;
    // 5-comparator sorting network on (distance, child); misses sort last
#define TPT_CSWAP(a, b)                                                                                                \
    if (tn[b] < tn[a]) {                                                                                               \
        float tt = tn[a];                                                                                              \
        tn[a] = tn[b];                                                                                                 \
        tn[b] = tt;                                                                                                    \
        int32_t cc = ch[a];                                                                                            \
        ch[a] = ch[b];                                                                                                 \
        ch[b] = cc;                                                                                                    \
    }
    TPT_CSWAP(0, 1)
    TPT_CSWAP(2, 3)
    TPT_CSWAP(0, 2)
    TPT_CSWAP(1, 3)
    TPT_CSWAP(1, 2)
#undef TPT_CSWAP
    if (nh > 3) stk.set(sp++, ch[3]);
    if (nh > 2) stk.set(sp++, ch[2]);
    if (nh > 1) stk.set(sp++, ch[1]);
    nearc = ch[0];
    return true;
}

// One traversal step: visit node `cur` (inner node: descend to the nearest hit
// child, push the others; leaf: test its triangles), then pop if needed.
// Returns false once the traversal is over.
template <bool ANY, bool COUNT, int NL, class Stk>
__device__ __forceinline__ bool trav_step(const DScene &S, V3 o, V3 d, const RayPre &rp, const AlphaCtx &ac, HitRec &best,
                                          int32_t &cur, int &sp, Stk &stk, uint32_t &nodes, uint32_t &tris, int cache) {
    if (cur >= 0) {
        if (COUNT) nodes++;
        int32_t nearc;
        if (visit_inner<NL>(S, cur, rp, cull_bound(rp, best.t), nearc, sp, stk, cache)) {
            cur = nearc;
            return true;
        }
    } else {
        bool done = false;
        leaf_tests(S, cur, o, d, ac, best, ANY, tris, COUNT, done);
        if (ANY && done) return false;
    }
    if (sp == 0) return false;
    cur = stk.get(--sp);
    return true;
}

// Closest hit (ANY = false) or any hit (ANY = true) with t in [EPS, tmax).
// Ties on t resolve to the smaller global primitive id.
template <bool ANY, bool COUNT, int NL, bool LDS>
__device__ __forceinline__ HitRec trace(const DScene &S, V3 o, V3 d, float tmax, const AlphaCtx &ac, uint32_t &nodes,
                                        uint32_t &tris) {
    HitRec best{tmax, -1, 0.f, 0.f};
    if (analytic_tests<ANY>(S, o, d, ac, best)) return best;
    if (S.root == -1) return best;
    RayPre rp = ray_pre(o, d);
    TStack<LDS> stk;
    stk.init(0);
    int sp = 0;
    int32_t cur = S.root;
    while (trav_step<ANY, COUNT, NL>(S, o, d, rp, ac, best, cur, sp, stk, nodes, tris, 0)) {
    }
    return best;
}

// ---------------------------------------------------------------- wave-level compaction helpers
__device__ __forceinline__ int wave_append(bool pred, uint32_t *counter) {
    uint64_t mask = __ballot(pred);
    int lane = (int)__lane_id();
    int total = __popcll(mask);
    int prefix = __popcll(mask & ((1ull << lane) - 1ull));
    uint32_t base = 0;
    int leader = __ffsll((unsigned long long)mask) - 1;
    if (total > 0 && lane == leader) base = atomicAdd(counter, (uint32_t)total);
    base = __shfl(base, leader < 0 ? 0 : leader);
    return pred ? (int)base + prefix : -1;
}

__device__ __forceinline__ void wave_count(uint32_t v32, uint32_t *counter) {
    // 64-bit counter at `counter` (8-B aligned); reduce within the wave, one atomic per wave
    unsigned long long v = v32;
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off);
    if (__lane_id() == 0 && v) atomicAdd(reinterpret_cast<unsigned long long *>(counter), v);
}

// ---------------------------------------------------------------- persistent traversal
// Each wave keeps its 64 lanes busy: a lane that finishes its ray takes a new
// one from a global counter as soon as kRefillLanes lanes of the wave are idle
// (one atomic per refill for the whole wave), instead of idling until the
// slowest ray of the wave is done.  Ray sources / sinks:
// Streamed path-state records (each read or written once per launch) use
// non-temporal loads / stores, so they do not evict BVH nodes and triangles
// from the L2 while the traversal runs.
__device__ __forceinline__ float4 ld_stream(const float4 *p) {
    return from_v4f(__builtin_nontemporal_load(reinterpret_cast<const v4f *>(p)));
}
__device__ __forceinline__ void st_stream(float4 *p, float4 v) {
    __builtin_nontemporal_store(v4f{v.x, v.y, v.z, v.w}, reinterpret_cast<v4f *>(p));
}
template <class T> __device__ __forceinline__ T ld_stream(const T *p) { return __builtin_nontemporal_load(p); }

struct SrcQueue { // extend: ray queue of slot ids -> hit records by queue position
    const int32_t *queue;
    const float4 *org, *dir;
    const uint64_t *key;
    const int32_t *depth;
    float4 *hit;
    __device__ __forceinline__ void load(uint32_t i, V3 &o, V3 &d, float &tmax, AlphaCtx &ac) const {
        int slot = ld_stream(queue + i);
        o = mk4(ld_stream(org + slot));
        d = mk4(ld_stream(dir + slot));
        tmax = kINF;
        ac = AlphaCtx{ld_stream(key + slot), ld_stream(depth + slot), true};
    }
    __device__ __forceinline__ void store(uint32_t i, const HitRec &h) const {
        st_stream(hit + i, make_float4(h.t, __int_as_float(h.gid), h.u, h.v));
    }
};
struct SrcShadow { // delta-light shadow rays -> add the contribution when unoccluded
    const float4 *org, *dir, *val;
    const int32_t *slot;
    float4 *color;
    __device__ __forceinline__ void load(uint32_t i, V3 &o, V3 &d, float &tmax, AlphaCtx &ac) const {
        float4 o4 = ld_stream(org + i);
        o = mk4(o4);
        d = mk4(ld_stream(dir + i));
        tmax = o4.w;
        ac = AlphaCtx{0, 0, false};
    }
    __device__ __forceinline__ void store(uint32_t i, const HitRec &h) const {
        if (h.gid < 0) {
            int s = slot[i];
            float4 c = color[s], v = val[i];
            color[s] = make_float4(c.x + v.x, c.y + v.y, c.z + v.z, 0.f);
        }
    }
};
struct SrcRays { // ray-query API: (org, tmin, dir, tmax) records -> prim id + t, or occlusion byte
    const float4 *rays;
    int32_t *prim;
    float *t;
    uint8_t *occ;
    __device__ __forceinline__ void load(uint32_t i, V3 &o, V3 &d, float &tmax, AlphaCtx &ac) const {
        float4 a = rays[2 * (size_t)i], b = rays[2 * (size_t)i + 1];
        o = mk4(a);
        d = mk4(b);
        tmax = b.w < kINF ? b.w : kINF;
        ac = AlphaCtx{0, 0, false};
    }
    __device__ __forceinline__ void store(uint32_t i, const HitRec &h) const {
        bool hit = h.gid >= 0;
        if (occ) {
            occ[i] = hit ? 1 : 0;
        } else {
            prim[i] = hit ? h.gid : -1;
            t[i] = hit ? h.t : kINF;
        }
    }
};

static constexpr int kRefillLanes = 16;
static constexpr int kLeafBatch = 16;
static constexpr int32_t kNoNode = -1; // never a leaf reference (those stay above 0xC0000000)

// Inner-node visit only: `cur` becomes the nearest hit child (inner node or
// leaf), else the next stack entry, else kNoNode.
template <bool COUNT, int NL, class Stk>
__device__ __forceinline__ void inner_visit(const DScene &S, const RayPre &rp, float tmax, int32_t &cur, int &sp,
                                            Stk &stk, uint32_t &nodes, int cache) {
    if (COUNT) nodes++;
    int32_t nearc;
    if (visit_inner<NL>(S, cur, rp, cull_bound(rp, tmax), nearc, sp, stk, cache)) {
        cur = nearc;
        return;
    }
    cur = sp > 0 ? stk.get(--sp) : kNoNode;
}

template <bool ANY, bool COUNT, int NL, bool POSTPONE, bool LDS, class Src>
__device__ __forceinline__ void persistent_trace(const DScene &S, const Src &src, uint32_t n, uint32_t *fetch,
                                                 uint32_t &nodes, uint32_t &tris) {
    int32_t leaf = kNoNode; // parked leaf of this lane
    // stage the top of the tree (quantized layout) in LDS: every ray starts there
    const int cache = NL == kNodeBvh4Q ? S.lds_nodes : 0;
    if (cache > 0) {
        lds_f4 *dst = (lds_f4 *)tpt_lds_stack;
        for (int i = threadIdx.x; i < 4 * cache; i += blockDim.x) {
            float4 v = S.nodes[i];
            dst[i] = v4f{v.x, v.y, v.z, v.w};
        }
        __syncthreads();
    }
    TStack<LDS> stk;
    stk.init(cache);
    const int lane = (int)__lane_id();
    const uint64_t lower = (1ull << lane) - 1ull;
    int64_t ray = -1; // ray index of this lane, -1 when idle
    bool exhausted = false;
    V3 o{0, 0, 0}, d{0, 0, 0};
    RayPre rp{};
    HitRec best{kINF, -1, 0.f, 0.f};
    AlphaCtx ac{0, 0, false};
    int32_t cur = -1;
    int sp = 0;
    while (true) {
        // refill idle lanes
        uint64_t need = __ballot(ray < 0 && !exhausted);
        uint64_t act = __ballot(ray >= 0);
        int nneed = __popcll(need);
        if (nneed > 0 && (nneed >= kRefillLanes || act == 0)) {
            int leader = __ffsll((unsigned long long)need) - 1;
            uint32_t base = 0;
            if (lane == leader) base = atomicAdd(fetch, (uint32_t)nneed);
            base = __shfl(base, leader);
            if ((uint64_t)base + (uint64_t)nneed >= (uint64_t)n) exhausted = true;
            if ((need >> lane) & 1ull) {
                uint32_t idx = base + (uint32_t)__popcll(need & lower);
                if (idx < n) {
                    float tmax;
                    src.load(idx, o, d, tmax, ac);
                    best = HitRec{tmax, -1, 0.f, 0.f};
                    if (analytic_tests<ANY>(S, o, d, ac, best) || S.root == -1) {
                        src.store(idx, best);
                    } else {
                        ray = idx;
                        rp = ray_pre(o, d);
                        cur = S.root;
                        sp = 0;
                    }
                }
            }
            act = __ballot(ray >= 0);
        }
        if (act == 0) {
            if (exhausted) break;
            continue;
        }
        // traverse until enough lanes are idle to make a refill worthwhile
        if (POSTPONE) {
            // Leaves are parked (one per lane) while the lane keeps descending inner
            // nodes; parked leaves are tested together once kLeafBatch lanes hold one
            // or no lane has inner work left, so most wave iterations run only the
            // inner-node code instead of inner + triangle code.
            while (true) {
                if (ray >= 0 && cur >= 0) inner_visit<COUNT, NL>(S, rp, best.t, cur, sp, stk, nodes, cache);
                if (ray >= 0 && leaf == kNoNode && cur != kNoNode && cur < 0) {
                    leaf = cur;
                    cur = sp > 0 ? stk.get(--sp) : kNoNode;
                }
                uint64_t lm = __ballot(ray >= 0 && leaf != kNoNode);
                uint64_t im = __ballot(ray >= 0 && cur >= 0);
                if (lm != 0 && (__popcll(lm) >= kLeafBatch || im == 0)) {
                    if (ray >= 0 && leaf != kNoNode) {
                        bool done = false;
                        leaf_tests(S, leaf, o, d, ac, best, ANY, tris, COUNT, done);
                        leaf = kNoNode;
                        if (ANY && done) {
                            cur = kNoNode;
                            sp = 0;
                        }
                    }
                }
                if (ray >= 0 && cur == kNoNode && leaf == kNoNode) {
                    src.store((uint32_t)ray, best);
                    ray = -1;
                }
                int active = __popcll(__ballot(ray >= 0));
                if (active == 0 || (!exhausted && 64 - active >= kRefillLanes)) break;
            }
        } else {
            while (true) {
                if (ray >= 0) {
                    if (!trav_step<ANY, COUNT, NL>(S, o, d, rp, ac, best, cur, sp, stk, nodes, tris, cache)) {
                        src.store((uint32_t)ray, best);
                        ray = -1;
                    }
                }
                int active = __popcll(__ballot(ray >= 0));
                if (active == 0 || (!exhausted && 64 - active >= kRefillLanes)) break;
            }
        }
    }
}

// Extra attributes of the persistent traversal kernels (occupancy-tuning builds
// define it, e.g. as __attribute__((amdgpu_waves_per_eu(8, 8)))).
#ifndef TPT_TRAVERSAL_ATTR
#define TPT_TRAVERSAL_ATTR
#endif

template <bool COUNT, int NL, bool POST, bool LDS>
__global__ TPT_TRAVERSAL_ATTR void k_extend_pt(DScene S, DPaths P, const int32_t *queue, int qsel, int nsel_next) {
    uint32_t n = P.counters[qsel];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        P.counters[nsel_next] = 0u; // next ray queue
        P.counters[2] = 0u;         // shadow queue
        P.counters[13] = 0u;        // shadow fetch counter
        if (COUNT) atomicAdd(reinterpret_cast<unsigned long long *>(P.counters + 8), (unsigned long long)n);
    }
    uint32_t nodes = 0, tris = 0;
    SrcQueue src{queue, P.org, P.dir, P.key, P.depth, P.hit};
    persistent_trace<false, COUNT, NL, POST, LDS>(S, src, n, P.counters + 12, nodes, tris);
    if (COUNT) {
        wave_count(nodes, P.counters + 4);
        wave_count(tris, P.counters + 6);
    }
}

template <bool COUNT, int NL, bool POST, bool LDS>
__global__ TPT_TRAVERSAL_ATTR void k_shadow_pt(DScene S, DPaths P) {
    uint32_t n = P.counters[2];
    if (COUNT && blockIdx.x == 0 && threadIdx.x == 0)
        atomicAdd(reinterpret_cast<unsigned long long *>(P.counters + 10), (unsigned long long)n);
    uint32_t nodes = 0, tris = 0;
    SrcShadow src{P.shadow_org, P.shadow_dir, P.shadow_val, P.shadow_slot, P.color};
    persistent_trace<true, false, NL, POST, LDS>(S, src, n, P.counters + 13, nodes, tris);
}

template <bool ANY, bool COUNT, int NL, bool POST, bool LDS>
__global__ TPT_TRAVERSAL_ATTR void k_trace_pt(DScene S, SrcRays src, uint32_t n, uint32_t *counters) {
    uint32_t nodes = 0, tris = 0;
    persistent_trace<ANY, COUNT, NL, POST, LDS>(S, src, n, counters + 8, nodes, tris);
    if (COUNT) {
        wave_count(nodes, counters + 0);
        wave_count(tris, counters + 2);
    }
}

// ---------------------------------------------------------------- generate
__global__ void k_generate(DScene S, DRender R, DPaths P, int wave_base, int n_paths, uint64_t base_mixed) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j == 0) {
        P.counters[0] = (uint32_t)n_paths;
        P.counters[12] = 0u; // extend fetch counter
    }
    if (j >= n_paths) return;
    int pix = j % R.pixels;
    int s = wave_base + j / R.pixels;
    int W = S.cam.width;
    int x = pix % W, y = pix / W;
    uint64_t key = stream_seed(base_mixed, (uint64_t)pix, (uint64_t)s);
    Rng rng{rng_state_from_seed(key)};
    // PixelSampler sub-pixel origin (path_tracer.cpp:8-58): repeated additions of 1/b
    int subpix = s / R.nsub;
    int sx = subpix % R.block, sy = subpix / R.block;
    float w = 1.f / (float)R.block;
    float px = (float)x - .5f, py = (float)y - .5f;
    for (int i = 0; i < sx; i++) px += w;
    for (int i = 0; i < sy; i++) py += w;
    px = px + rng.random(w);
    py = py + rng.random(w);
    // Camera::shoot_ray (camera.cpp:23-34)
    const DCamera &c = S.cam;
    V3 offset{0.f, 0.f, 0.f};
    if (c.aperture >= kEPS) {
        V2 uo = rng.rand_in_disk(c.aperture);
        offset = add(scl(uo.x, mk(c.u)), scl(uo.y, mk(c.v)));
    }
    float a = px / (float)c.width - .5f, b = py / (float)c.height - .5f;
    V3 dir = sub(add(add(scl(a, mk(c.right)), scl(b, mk(c.up))), mk(c.direct)), offset);
    dir = nrm(dir);
    V3 org = add(mk(c.center), offset);
    P.org[j] = make_float4(org.x, org.y, org.z, 0.f);
    P.dir[j] = make_float4(dir.x, dir.y, dir.z, 0.f);
    P.decay[j] = make_float4(1.f, 1.f, 1.f, 0.f);
    P.color[j] = make_float4(0.f, 0.f, 0.f, 0.f);
    P.rng[j] = rng.x;
    P.key[j] = key;
    P.depth[j] = 0;
    P.queue_a[j] = j;
}

// ---------------------------------------------------------------- extend
template <bool COUNT, int NL>
__global__ void k_extend(DScene S, DPaths P, const int32_t *queue, int qsel, int nsel_next) {
    uint32_t n = P.counters[qsel];
    int tid = blockIdx.x * blockDim.x + threadIdx.x;
    if (tid == 0) {
        P.counters[nsel_next] = 0u; // next ray queue
        P.counters[2] = 0u;         // shadow queue
        if (COUNT) atomicAdd(reinterpret_cast<unsigned long long *>(P.counters + 8), (unsigned long long)n);
    }
    uint32_t nodes = 0, tris = 0;
    for (uint32_t i = (uint32_t)tid; i < n; i += gridDim.x * blockDim.x) {
        int slot = queue[i];
        float4 o4 = P.org[slot], d4 = P.dir[slot];
        AlphaCtx ac{P.key[slot], P.depth[slot], true};
        HitRec h = trace<false, COUNT, NL, false>(S, mk4(o4), mk4(d4), kINF, ac, nodes, tris);
        P.hit[i] = make_float4(h.t, __int_as_float(h.gid), h.u, h.v);
    }
    if (COUNT) {
        wave_count(nodes, P.counters + 4);
        wave_count(tris, P.counters + 6);
    }
}

// ---------------------------------------------------------------- shade
__device__ __forceinline__ void surface(const DScene &S, V3 o, V3 d, const HitRec &h, SurfHit &sh, int &mat) {
    int32_t loc = S.gid_local[h.gid];
    if (loc < 0) {
        const DAnalytic &p = S.analytic[-loc - 1];
        mat = p.material;
        if (p.type == A_SPHERE) { // object.cpp:44-59
            V3 pos = at(o, d, h.t);
            V3 n = nrm(sub(pos, mk(p.center)));
            float u = (atan2f(n.y, n.x) + kPI) * (.5f * kINV_PI);
            float v = acosf(n.z) * kINV_PI;
            bool into = h.u != 0.f;
            if (!into) n = neg(n);
            sh = SurfHit{n, n, V2{u, v}, into};
        } else { // object.cpp:123-126, 170-174
            V2 uv;
            plane_t(p, o, d, kINF, uv);
            V3 n = mk(p.w);
            V3 hn = same_side(n, d) ? neg(n) : n;
            sh = SurfHit{hn, hn, uv, true};
        }
        return;
    }
    // triangle (object.cpp:266-331)
    int tri = loc;
    int4 i0 = S.tri_info[2 * tri], i1 = S.tri_info[2 * tri + 1];
    mat = i0.x;
    float u = h.u, v = h.v;
    float4 r1 = S.tris[3 * tri + 1], r2 = S.tris[3 * tri + 2];
    V3 ab = mk4(r1), ac = mk4(r2);
    V2 uv;
    const float *tc = S.texcoords;
    if (i1.x >= 0) {
        float w0 = 1.f - u - v;
        uv = {(w0 * tc[2 * i1.x] + u * tc[2 * i1.y]) + v * tc[2 * i1.z],
              (w0 * tc[2 * i1.x + 1] + u * tc[2 * i1.y + 1]) + v * tc[2 * i1.z + 1]};
    } else {
        uv = {u, v};
    }
    V3 n;
    if (i0.y >= 0) {
        const float *nm = S.normals;
        float w0 = 1.f - u - v;
        V3 na = mk(nm + 3 * i0.y), nb = mk(nm + 3 * i0.z), nc = mk(nm + 3 * i0.w);
        n = nrm(add(add(scl(w0, na), scl(u, nb)), scl(v, nc)));
    } else {
        n = nrm(cross(ac, ab));
    }
    bool into = !same_side(n, d);
    if (!into) n = neg(n);
    V3 sn = n;
    const DMaterial &m = S.materials[mat];
    if (m.bump >= 0) {
        V2 dab, dac;
        if (i1.x >= 0) {
            dab = {tc[2 * i1.y] - tc[2 * i1.x], tc[2 * i1.y + 1] - tc[2 * i1.x + 1]};
            dac = {tc[2 * i1.z] - tc[2 * i1.x], tc[2 * i1.z + 1] - tc[2 * i1.x + 1]};
        } else {
            dab = {1.f, 0.f};
            dac = {0.f, 1.f};
        }
        float inv_det = 1.f / (dab.x * dac.y - dac.x * dab.y);
        if (isfinite(inv_det)) {
            V3 dpdu = scr(sub(scl(dac.y, ab), scl(dab.y, ac)), inv_det);
            V3 dpdv = scr(add(scl(-dac.x, ab), scl(dab.x, ac)), inv_det);
            sn = bump_normal(S, m.bump, uv, sn, dpdu, dpdv);
        }
    }
    sh = SurfHit{n, sn, uv, into};
}

__global__ void k_shade(DScene S, DRender R, DPaths P, const int32_t *queue, int32_t *next, int qsel, int nsel) {
    uint32_t n = P.counters[qsel];
    if (blockIdx.x == 0 && threadIdx.x == 0) P.counters[12] = 0u; // fetch counter of the next extend
    const float LIGHT_PROB = 1.f / 5.f; // cpu/path_tracer.cpp:104
    for (uint32_t base = blockIdx.x * blockDim.x; base < n; base += gridDim.x * blockDim.x) {
        uint32_t i = base + threadIdx.x;
        bool valid = i < n;
        bool alive = false;
        bool emit_shadow = false;
        int slot = 0;
        V3 s_org{0, 0, 0}, s_dir{0, 0, 0}, s_val{0, 0, 0};
        float s_tmax = kINF;
        if (valid) {
            slot = queue[i];
            float4 h4 = P.hit[i];
            HitRec h{h4.x, __float_as_int(h4.y), h4.z, h4.w};
            V3 o = mk4(P.org[slot]), d = mk4(P.dir[slot]);
            V3 decay = mk4(P.decay[slot]);
            V3 color = mk4(P.color[slot]);
            int depth = P.depth[slot];
            Rng rng{P.rng[slot]};
            if (!(h.t < kINF)) {
                color = add(color, cmul(decay, background(S, d)));
            } else {
                SurfHit sh;
                int mi;
                surface(S, o, d, h, sh, mi);
                const DMaterial &m = S.materials[mi];
                color = add(color, cmul(decay, color_at(S, m.emission, sh.uv)));
                if (depth < R.max_depth) {
                    V3 hp = at(o, d, h.t);
                    float pdf_bxdf;
                    bool spec;
                    int sel = sample_bxdf(m, rng, pdf_bxdf, spec);
                    V3 out, albedo;
                    if (!spec) {
                        if (S.num_delta > 0) {
                            const DDelta &dl = S.delta[rng.rand_int(S.num_delta)];
                            float tmax;
                            V3 tl = delta_sample(dl, hp, rng, tmax);
                            if (!R.shadow_bounded) tmax = kINF;
                            V3 la = bxdf_f(S, m, sel, d, sh, tl);
                            float pdf_light = 1.f / (float)S.num_delta;
                            V3 contrib = dvs(cmul(cmul(decay, la), mk(dl.emission)), pdf_light);
                            if (contrib.x != 0.f || contrib.y != 0.f || contrib.z != 0.f) {
                                emit_shadow = true;
                                s_org = hp;
                                s_dir = tl;
                                s_val = contrib;
                                s_tmax = tmax;
                            }
                        }
                        float pdf_val;
                        if (S.num_lights > 0) {
                            if (rng.random() < LIGHT_PROB) {
                                out = lights_sample(S, hp, rng);
                            } else {
                                out = bxdf_sample(m, sel, d, sh, rng);
                            }
                            pdf_val = (1.f - LIGHT_PROB) * bxdf_pdf(m, sel, d, sh, out) +
                                      LIGHT_PROB * lights_pdf(S, hp, out);
                        } else {
                            out = bxdf_sample(m, sel, d, sh, rng);
                            pdf_val = bxdf_pdf(m, sel, d, sh, out);
                        }
                        albedo = scr(bxdf_f(S, m, sel, d, sh, out), 1.f / ((pdf_val + kEPS) * pdf_bxdf));
                    } else {
                        out = bxdf_sample(m, sel, d, sh, rng);
                        albedo = scr(bxdf_f(S, m, sel, d, sh, out), 1.f / pdf_bxdf);
                    }
                    decay = cmul(decay, albedo);
                    alive = !is_zero5(decay);
                    if (alive && R.rr_depth >= 0 && depth >= R.rr_depth) {
                        float mx = fmaxf(decay.x, fmaxf(decay.y, decay.z));
                        float q = fminf(mx, 0.95f);
                        if (rng.random() >= q) {
                            alive = false;
                        } else {
                            decay = dvs(decay, q);
                        }
                    }
                    if (alive) {
                        P.org[slot] = make_float4(hp.x, hp.y, hp.z, 0.f);
                        P.dir[slot] = make_float4(out.x, out.y, out.z, 0.f);
                        P.decay[slot] = make_float4(decay.x, decay.y, decay.z, 0.f);
                        P.depth[slot] = depth + 1;
                    }
                }
            }
            P.color[slot] = make_float4(color.x, color.y, color.z, 0.f);
            P.rng[slot] = rng.x;
        }
        int pos = wave_append(alive, P.counters + nsel);
        if (alive) next[pos] = slot;
        int spos = wave_append(emit_shadow, P.counters + 2);
        if (emit_shadow) {
            P.shadow_slot[spos] = slot;
            P.shadow_org[spos] = make_float4(s_org.x, s_org.y, s_org.z, s_tmax);
            P.shadow_dir[spos] = make_float4(s_dir.x, s_dir.y, s_dir.z, 0.f);
            P.shadow_val[spos] = make_float4(s_val.x, s_val.y, s_val.z, 0.f);
        }
    }
}

// ---------------------------------------------------------------- shadow
template <bool COUNT, int NL>
__global__ void k_shadow(DScene S, DPaths P) {
    uint32_t n = P.counters[2];
    if (COUNT && blockIdx.x == 0 && threadIdx.x == 0)
        atomicAdd(reinterpret_cast<unsigned long long *>(P.counters + 10), (unsigned long long)n);
    uint32_t nodes = 0, tris = 0;
    AlphaCtx ac{0, 0, false};
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        float4 o4 = P.shadow_org[i], d4 = P.shadow_dir[i];
        HitRec h = trace<true, false, NL, false>(S, mk4(o4), mk4(d4), o4.w, ac, nodes, tris);
        if (h.gid < 0) {
            int slot = P.shadow_slot[i];
            float4 c = P.color[slot], v = P.shadow_val[i];
            P.color[slot] = make_float4(c.x + v.x, c.y + v.y, c.z + v.z, 0.f);
        }
    }
}

// ---------------------------------------------------------------- accumulate / finalize
__global__ void k_accumulate(DRender R, DPaths P, int wave_base, int k_active, int first_wave) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= R.pixels) return;
    float sf[3], pf[3];
    double sd[3], pd[3];
    for (int c = 0; c < 3; c++) {
        sf[c] = first_wave ? 0.f : P.sub_f[3 * p + c];
        pf[c] = first_wave ? 0.f : P.pix_f[3 * p + c];
        sd[c] = (first_wave || !R.accum_fp64) ? 0.0 : P.sub_d[3 * p + c];
        pd[c] = (first_wave || !R.accum_fp64) ? 0.0 : P.pix_d[3 * p + c];
    }
    for (int j = 0; j < k_active; j++) {
        int s = wave_base + j;
        float4 L = P.color[(size_t)j * R.pixels + p];
        if (R.accum_fp64) {
            sd[0] += L.x;
            sd[1] += L.y;
            sd[2] += L.z;
        } else {
            sf[0] = sf[0] + L.x;
            sf[1] = sf[1] + L.y;
            sf[2] = sf[2] + L.z;
        }
        if ((s + 1) % R.nsub == 0) { // sub-pixel complete (path_tracer.cpp:91)
            for (int c = 0; c < 3; c++) {
                if (R.accum_fp64) {
                    if (R.output == 1) pd[c] += sd[c];
                    else pf[c] = pf[c] + clip1((float)(sd[c] / R.nsub));
                    sd[c] = 0.0;
                } else {
                    if (R.output == 1) pf[c] = pf[c] + sf[c];
                    else pf[c] = pf[c] + clip1(sf[c] / (float)R.nsub);
                    sf[c] = 0.f;
                }
            }
        }
    }
    for (int c = 0; c < 3; c++) {
        P.sub_f[3 * p + c] = sf[c];
        P.pix_f[3 * p + c] = pf[c];
        if (R.accum_fp64) {
            P.sub_d[3 * p + c] = sd[c];
            P.pix_d[3 * p + c] = pd[c];
        }
    }
}

__global__ void k_finalize(DRender R, DPaths P, float *out) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= R.pixels) return;
    int bb = R.block * R.block;
    for (int c = 0; c < 3; c++) {
        float v;
        if (R.output == 1) {
            if (R.accum_fp64) v = (float)(P.pix_d[3 * p + c] / ((double)R.nsub * (double)bb));
            else v = P.pix_f[3 * p + c] / (float)(R.nsub * bb);
        } else {
            v = clip1(P.pix_f[3 * p + c] / (float)bb);
        }
        out[3 * p + c] = v;
    }
}

// ---------------------------------------------------------------- ray queries (parity / bench)
template <bool COUNT, int NL>
__global__ void k_trace_closest(DScene S, const float4 *rays, uint64_t n, int32_t *prim, float *tout,
                                uint32_t *counters) {
    uint32_t nodes = 0, tris = 0;
    AlphaCtx ac{0, 0, false};
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        float4 a = rays[2 * i], b = rays[2 * i + 1];
        float tmax = b.w < kINF ? b.w : kINF;
        HitRec h = trace<false, COUNT, NL, false>(S, mk4(a), mk4(b), tmax, ac, nodes, tris);
        bool hit = h.gid >= 0;
        prim[i] = hit ? h.gid : -1;
        tout[i] = hit ? h.t : kINF;
    }
    if (COUNT) {
        wave_count(nodes, counters + 0);
        wave_count(tris, counters + 2);
    }
}

template <int NL>
__global__ void k_trace_occluded(DScene S, const float4 *rays, uint64_t n, uint8_t *occ) {
    uint32_t nodes = 0, tris = 0;
    AlphaCtx ac{0, 0, false};
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        float4 a = rays[2 * i], b = rays[2 * i + 1];
        float tmax = b.w < kINF ? b.w : kINF;
        HitRec h = trace<true, false, NL, false>(S, mk4(a), mk4(b), tmax, ac, nodes, tris);
        occ[i] = h.gid >= 0 ? 1 : 0;
    }
}

// ---------------------------------------------------------------- host-visible launchers
static inline size_t lds_bytes(int block) { return (size_t)kLdsStack * (size_t)block * sizeof(int32_t); }

// Persistent kernels launch exactly the resident grid: num_cu x (blocks per CU
// the occupancy calculator allows for this kernel, block size and LDS).
static int resident_grid(const DScene &S, const void *kernel, int block, size_t shared, int wanted) {
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, block, shared) != hipSuccess || per_cu < 1)
        per_cu = 1;
    int g = S.num_cu * per_cu;
    return wanted < g ? (wanted < 1 ? 1 : wanted) : g;
}

void launch_generate(const DScene &S, const DRender &R, const DPaths &P, int wave_base, int n_paths,
                     uint64_t base_mixed, int block, hipStream_t st) {
    int grid = (n_paths + block - 1) / block;
    if (grid < 1) grid = 1;
    hipLaunchKernelGGL(k_generate, dim3(grid), dim3(block), 0, st, S, R, P, wave_base, n_paths, base_mixed);
}
// Compile-time dispatch on the node layout and on boolean kernel variants.
template <int V> using IntC = std::integral_constant<int, V>;
template <class F> static void by_layout(int nl, F &&f) {
    if (nl == kNodeBvh2) f(IntC<kNodeBvh2>{});
    else if (nl == kNodeBvh4) f(IntC<kNodeBvh4>{});
    else f(IntC<kNodeBvh4Q>{});
}
template <class F> static void by_bool(bool b, F &&f) {
    if (b) f(std::true_type{});
    else f(std::false_type{});
}
// Persistent-kernel variants: parked leaves (S.persistent == 2) and LDS stack.
template <class F> static void by_variant(const DScene &S, F &&f) {
    by_bool(S.persistent == 2, [&](auto Q) { by_bool(S.lds_stack != 0, [&](auto L) { f(Q, L); }); });
}
// Dynamic LDS of a persistent traversal launch: cached nodes + the LDS stack.
static size_t trav_shared(const DScene &S, int nl, bool lds_stack) {
    size_t nodes = nl == kNodeBvh4Q ? (size_t)S.lds_nodes * 64 : 0;
    return nodes + (lds_stack ? lds_bytes(S.trav_block) : 0);
}

void launch_extend(const DScene &S, const DPaths &P, const int32_t *queue, int qsel, int nsel, int grid, int block,
                   bool count, hipStream_t st) {
    by_layout(S.layout, [&](auto L) {
        by_bool(count, [&](auto C) {
            constexpr int NL = decltype(L)::value;
            constexpr bool CNT = decltype(C)::value;
            if (!S.persistent) {
                hipLaunchKernelGGL((k_extend<CNT, NL>), dim3(grid), dim3(block), 0, st, S, P, queue, qsel, nsel);
                return;
            }
            by_variant(S, [&](auto Q, auto D) {
                constexpr bool POST = decltype(Q)::value, LDS = decltype(D)::value;
                const int tb = S.trav_block;
                size_t sh = trav_shared(S, NL, LDS);
                int g = resident_grid(S, (const void *)(k_extend_pt<CNT, NL, POST, LDS>), tb, sh, grid);
                hipLaunchKernelGGL((k_extend_pt<CNT, NL, POST, LDS>), dim3(g), dim3(tb), sh, st, S, P, queue, qsel,
                                   nsel);
            });
        });
    });
}
void launch_shade(const DScene &S, const DRender &R, const DPaths &P, const int32_t *queue, int32_t *next, int qsel,
                  int nsel, int grid, int block, hipStream_t st) {
    hipLaunchKernelGGL(k_shade, dim3(grid), dim3(block), 0, st, S, R, P, queue, next, qsel, nsel);
}
void launch_shadow(const DScene &S, const DPaths &P, int grid, int block, bool count, hipStream_t st) {
    by_layout(S.layout, [&](auto L) {
        by_bool(count, [&](auto C) {
            constexpr int NL = decltype(L)::value;
            constexpr bool CNT = decltype(C)::value;
            if (!S.persistent) {
                hipLaunchKernelGGL((k_shadow<CNT, NL>), dim3(grid), dim3(block), 0, st, S, P);
                return;
            }
            by_variant(S, [&](auto Q, auto D) {
                constexpr bool POST = decltype(Q)::value, LDS = decltype(D)::value;
                const int tb = S.trav_block;
                size_t sh = trav_shared(S, NL, LDS);
                int g = resident_grid(S, (const void *)(k_shadow_pt<CNT, NL, POST, LDS>), tb, sh, grid);
                hipLaunchKernelGGL((k_shadow_pt<CNT, NL, POST, LDS>), dim3(g), dim3(tb), sh, st, S, P);
            });
        });
    });
}
void launch_accumulate(const DRender &R, const DPaths &P, int wave_base, int k_active, int first, hipStream_t st) {
    int grid = (R.pixels + 255) / 256;
    hipLaunchKernelGGL(k_accumulate, dim3(grid), dim3(256), 0, st, R, P, wave_base, k_active, first);
}
void launch_finalize(const DRender &R, const DPaths &P, float *out, hipStream_t st) {
    int grid = (R.pixels + 255) / 256;
    hipLaunchKernelGGL(k_finalize, dim3(grid), dim3(256), 0, st, R, P, out);
}
void launch_trace_closest(const DScene &S, const float4 *rays, uint64_t n, int32_t *prim, float *t,
                          uint32_t *counters, bool count, int grid, int block, hipStream_t st) {
    by_layout(S.layout, [&](auto L) {
        by_bool(count, [&](auto C) {
            constexpr int NL = decltype(L)::value;
            constexpr bool CNT = decltype(C)::value;
            if (!S.persistent) {
                hipLaunchKernelGGL((k_trace_closest<CNT, NL>), dim3(grid), dim3(block), 0, st, S, rays, n, prim, t,
                                   counters);
                return;
            }
            SrcRays src{rays, prim, t, nullptr};
            by_variant(S, [&](auto Q, auto D) {
                constexpr bool POST = decltype(Q)::value, LDS = decltype(D)::value;
                const int tb = S.trav_block;
                size_t sh = trav_shared(S, NL, LDS);
                int g = resident_grid(S, (const void *)(k_trace_pt<false, CNT, NL, POST, LDS>), tb, sh, grid);
                hipLaunchKernelGGL((k_trace_pt<false, CNT, NL, POST, LDS>), dim3(g), dim3(tb), sh, st, S, src,
                                   (uint32_t)n, counters);
            });
        });
    });
}
void launch_trace_occluded(const DScene &S, const float4 *rays, uint64_t n, uint8_t *occ, uint32_t *counters, int grid,
                           int block, hipStream_t st) {
    by_layout(S.layout, [&](auto L) {
        constexpr int NL = decltype(L)::value;
        if (!S.persistent) {
            hipLaunchKernelGGL((k_trace_occluded<NL>), dim3(grid), dim3(block), 0, st, S, rays, n, occ);
            return;
        }
        SrcRays src{rays, nullptr, nullptr, occ};
        by_variant(S, [&](auto Q, auto D) {
            constexpr bool POST = decltype(Q)::value, LDS = decltype(D)::value;
            const int tb = S.trav_block;
            size_t sh = trav_shared(S, NL, LDS);
            int g = resident_grid(S, (const void *)(k_trace_pt<true, false, NL, POST, LDS>), tb, sh, grid);
            hipLaunchKernelGGL((k_trace_pt<true, false, NL, POST, LDS>), dim3(g), dim3(tb), sh, st, S, src,
                               (uint32_t)n, counters);
        });
    });
}

} // namespace tpt
