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
#include "nrc.h"

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
    float u, v;  // triangle barycentrics; analytic: u = into
    int32_t loc; // triangle id, or -(analytic index + 1)
};

// alpha test of a triangle candidate (object.cpp:273-287)
// Triangle id of a triangle location (DScene::loc_bits: the material index may sit above it).
__device__ __forceinline__ int32_t loc_tri(const DScene &S, int32_t loc) {
    return S.loc_bits ? loc & (int32_t)((1u << S.loc_bits) - 1u) : loc;
}

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
// [stack: kLdsRow * blockDim.x entries when the LDS stack is on].
extern __shared__ int32_t tpt_lds_stack[];

// Traversal stack: with LDS on, the first kLdsStack entries of every lane live
// in LDS, one row of kLdsRow = kLdsStack + 1 words per lane ([lane][entry]): an
// entry's address is the lane's row plus 4 * sp (a shift and an add, no
// multiply), and the odd row length puts the 64 lanes' entries of one depth in
// 64 different banks.  Deeper entries (rare: the host bounds the worst case by
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
    lds_i32 *lds; // this lane's row
    int32_t spill[kStackSize];
    // `cached_nodes`: 64-B node records stored in front of the stack
    __device__ __forceinline__ void init(int cached_nodes) {
        lds = (lds_i32 *)tpt_lds_stack + 16 * cached_nodes + (int)threadIdx.x * kLdsRow;
    }
    __device__ __forceinline__ void set(int sp, int32_t v) {
        if (LDS && sp < kLdsStack) lds[sp] = v;
        else spill[sp] = v;
    }
    __device__ __forceinline__ int32_t get(int sp) const {
        if (LDS && sp < kLdsStack) return lds[sp];
        return spill[sp];
    }
    // Entries sp .. sp + 2 are in LDS: put3 may write all three unconditionally.
    __device__ __forceinline__ bool room3(int sp) const { return LDS && sp <= kLdsStack - 3; }
    __device__ __forceinline__ void put3(int sp, int32_t a, int32_t b, int32_t c) {
        lds[sp] = a;
        lds[sp + 1] = b;
        lds[sp + 2] = c;
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
                if (!tri_alpha_pass(S, loc_tri(S, __float_as_int(r1.w)), u, v, ac, gid)) continue;
            }
            best = HitRec{t, gid, u, v, __float_as_int(r1.w)};
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
        const DAnalytic p = load_invariant(S.analytic + i); // i is wave-uniform: scalar loads
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
            best = HitRec{t, p.gid, into ? 1.f : 0.f, 0.f, -(i + 1)};
            if (ANY) return true;
        }
    }
    return false;
}

// TPT_NODE_FMA_MIX=1 builds the node test on v_fma_mix_f32 (below): 5 fewer VALU
// instructions per node visit, but extend alone took 249 against 239 ms per room
// frame (1024x768x256, same images), so the byte conversions stay the default.
#ifndef TPT_NODE_FMA_MIX
#define TPT_NODE_FMA_MIX 0
#endif
// Bytes 0, 1 (lo) or 2, 3 (hi) of w as two f16 values 1024 + byte (bits 0x64, byte).
__device__ __forceinline__ uint32_t f16_pair_lo(uint32_t w) { return __builtin_amdgcn_perm(0x64646464u, w, 0x04010400u); }
__device__ __forceinline__ uint32_t f16_pair_hi(uint32_t w) { return __builtin_amdgcn_perm(0x64646464u, w, 0x04030402u); }
// fma(h, a, b) with h the f16 in the low (hi = false) or high half of w, exactly
// converted (v_fma_mix_f32: one rounding, as fmaf of the converted value).
__device__ __forceinline__ float fma_mix16(uint32_t w, bool hi, float a, float b) {
    float r;
    if (hi) asm("v_fma_mix_f32 %0, %1, %2, %3 op_sel:[1,0,0] op_sel_hi:[1,0,0]" : "=v"(r) : "v"(w), "v"(a), "v"(b));
    else asm("v_fma_mix_f32 %0, %1, %2, %3 op_sel_hi:[1,0,0]" : "=v"(r) : "v"(w), "v"(a), "v"(b));
    return r;
}

// Slab test on precomputed slab distances (t0 / t1 per axis, any order).
__device__ __forceinline__ bool slab_hit(float t0x, float t1x, float t0y, float t1y, float t0z, float t1z, float tw,
                                         float &tnear) {
    float n = fmaxf(fmaxf(fminf(t0x, t1x), fminf(t0y, t1y)), fmaxf(fminf(t0z, t1z), kEPS * 0.99f));
    float f = fminf(fminf(fmaxf(t0x, t1x), fmaxf(t0y, t1y)), fmaxf(t0z, t1z));
    tnear = n;
    return n <= fminf(f * kWiden, tw);
}

// Slab test on per-axis near / far distances already ordered by the ray's
// direction (the same values as slab_hit's min / max of the two planes).
__device__ __forceinline__ bool ordered_hit(float nx, float fx, float ny, float fy, float nz, float fz, float tw,
                                           float &tnear) {
    float n = fmaxf(fmaxf(nx, ny), fmaxf(nz, kEPS * 0.99f));
    float f = fminf(fminf(fx, fy), fz);
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
        // near / far planes by the direction's signs: with inv > 0 a slab distance
        // grows with q (a = scale * inv > 0, fma is monotone), so lo is the near
        // plane; one select per packed word orders all four children
        const bool sx = rp.inv.x < 0.f, sy = rp.inv.y < 0.f, sz = rp.inv.z < 0.f;
        const uint32_t nqx = sx ? qhx : qlx, fqx = sx ? qlx : qhx;
        const uint32_t nqy = sy ? qhy : qly, fqy = sy ? qly : qhy;
        const uint32_t nqz = sz ? qhz : qlz, fqz = sz ? qlz : qhz;
#if TPT_NODE_FMA_MIX
        // Each plane byte q becomes the f16 value 1024 + q (bits 0x64, q: v_perm_b32
        // builds two per word), which v_fma_mix_f32 reads as an FMA operand, so
        // fma(1024 + q, a, b - 1024 a) replaces the byte-to-float conversion and the
        // FMA (12 perms + 24 mixed FMAs instead of 24 conversions + 24 FMAs).  The
        // extra rounding of b - 1024 a is 2^-24 max(|b|, 1024 |a|): against the
        // half-quantum margin |a| / 2 that keeps the test conservative, 2^-14 |a|.
        const float cx = fmaf(-1024.f, ax, bx), cy = fmaf(-1024.f, ay, by), cz = fmaf(-1024.f, az, bz);
        const uint32_t pn[3][2] = {{f16_pair_lo(nqx), f16_pair_hi(nqx)}, {f16_pair_lo(nqy), f16_pair_hi(nqy)},
                                   {f16_pair_lo(nqz), f16_pair_hi(nqz)}};
        const uint32_t pf[3][2] = {{f16_pair_lo(fqx), f16_pair_hi(fqx)}, {f16_pair_lo(fqy), f16_pair_hi(fqy)},
                                   {f16_pair_lo(fqz), f16_pair_hi(fqz)}};
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int w = i >> 1;
            const bool hi = i & 1;
            hb[i] = ordered_hit(fma_mix16(pn[0][w], hi, ax, cx), fma_mix16(pf[0][w], hi, ax, cx),
                                fma_mix16(pn[1][w], hi, ay, cy), fma_mix16(pf[1][w], hi, ay, cy),
                                fma_mix16(pn[2][w], hi, az, cz), fma_mix16(pf[2][w], hi, az, cz), tw, tn[i]);
        }
#else
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int sh = 8 * i;
            hb[i] = ordered_hit(fmaf((float)((nqx >> sh) & 0xffu), ax, bx), fmaf((float)((fqx >> sh) & 0xffu), ax, bx),
                                fmaf((float)((nqy >> sh) & 0xffu), ay, by), fmaf((float)((fqy >> sh) & 0xffu), ay, by),
                                fmaf((float)((nqz >> sh) & 0xffu), az, bz), fmaf((float)((fqz >> sh) & 0xffu), az, bz),
                                tw, tn[i]);
        }
#endif
    }
    // empty slots are sentinel leaves (never hit), so no slot check; a hit's
    // distance is at most tw <= kINF, a miss sorts last as +inf
    if (!(hb[0] | hb[1] | hb[2] | hb[3])) return false;
    constexpr float kMiss = __builtin_huge_valf();
#pragma unroll
    for (int i = 0; i < 4; i++) tn[i] = hb[i] ? tn[i] : kMiss;
    // 5-comparator sorting network on (distance, child); partial orders with the
    // last one or two comparators dropped measured 1-2% slower (room 1024x768x256)
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
    const bool h1 = tn[1] < kMiss, h2 = tn[2] < kMiss, h3 = tn[3] < kMiss;
    if (stk.room3(sp)) {
        // branch-free push of the far hit children, farthest deepest; the words
        // above the new top are scratch
        stk.put3(sp, h3 ? ch[3] : (h2 ? ch[2] : ch[1]), h3 ? ch[2] : ch[1], ch[1]);
        sp += (int)h1 + (int)h2 + (int)h3;
    } else {
        if (h3) stk.set(sp++, ch[3]);
        if (h2) stk.set(sp++, ch[2]);
        if (h1) stk.set(sp++, ch[1]);
    }
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
// `cache`: nodes [0, cache) are read from the workgroup's LDS copy (stage_top_nodes).
template <bool ANY, bool COUNT, int NL, bool LDS>
__device__ __forceinline__ HitRec trace(const DScene &S, V3 o, V3 d, float tmax, const AlphaCtx &ac, uint32_t &nodes,
                                        uint32_t &tris, int cache) {
    HitRec best{tmax, -1, 0.f, 0.f};
    if (analytic_tests<ANY>(S, o, d, ac, best)) return best;
    if (S.root == -1) return best;
    RayPre rp = ray_pre(o, d);
    TStack<LDS> stk;
    stk.init(cache);
    int sp = 0;
    int32_t cur = S.root;
    while (trav_step<ANY, COUNT, NL>(S, o, d, rp, ac, best, cur, sp, stk, nodes, tris, cache)) {
    }
    return best;
}

// Copies nodes [0, cache) of the quantized layout (the top of the tree: the
// node array starts in BFS order) to the front of the workgroup's LDS, where
// every ray starts.  All threads of the workgroup call it (it synchronises).
__device__ __forceinline__ void stage_top_nodes(const DScene &S, int cache) {
    if (cache <= 0) return;
    lds_f4 *dst = (lds_f4 *)tpt_lds_stack;
    for (int i = threadIdx.x; i < 4 * cache; i += blockDim.x) {
        float4 v = S.nodes[i];
        dst[i] = v4f{v.x, v.y, v.z, v.w};
    }
    __syncthreads();
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
// one as soon as kRefillLanes lanes of the wave are idle, instead of idling
// until the slowest ray of the wave is done.  New rays come from the wave's own
// chunk of the queue (FetchLocal, the default kernels) or from a counter shared
// by the launch (FetchGlobal, one atomic per refill).  Ray sources / sinks:
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

// Path rays carry their closest analytic hit (put_ray): org.w = t, dir.w = the
// analytic index * 2 + into, or -1.  Sources with kAnalyticDone start the
// traversal from it; the others test the analytic primitives when a lane
// takes the ray (ray_start).
struct SrcQueue { // extend: ray queue of slot ids -> hit records by queue position
    static constexpr bool kAnalyticDone = true;
    const int32_t *queue;
    const float4 *org, *dir;
    const uint64_t *key;
    const int32_t *depth;
    float4 *hit;
    __device__ __forceinline__ void load_hit(const DScene &S, uint32_t i, V3 &o, V3 &d, AlphaCtx &ac,
                                             HitRec &best) const {
        int slot = ld_stream(queue + i);
        float4 o4 = ld_stream(org + slot), d4 = ld_stream(dir + slot);
        o = mk4(o4);
        d = mk4(d4);
        ac = AlphaCtx{ld_stream(key + slot), ld_stream(depth + slot), true};
        const int code = __float_as_int(d4.w);
        best = code < 0 ? HitRec{o4.w, -1, 0.f, 0.f, 0}
                        : HitRec{o4.w, S.analytic[code >> 1].gid, (float)(code & 1), 0.f, -((code >> 1) + 1)};
    }
    __device__ __forceinline__ void store(uint32_t i, const HitRec &h) const {
        st_stream(hit + i, make_float4(h.t, __int_as_float(h.loc), h.u, h.v));
    }
};
struct SrcShadow { // delta-light shadow rays -> add the contribution when unoccluded
    static constexpr bool kAnalyticDone = false;
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
    static constexpr bool kAnalyticDone = false;
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

// Refill once 24 of the 64 lanes are idle; test parked leaves once 16 lanes hold
// one (room 1024x768x256, chunked kernels: refill at 8 / 16 / 24 / 32 idle lanes
// gives 401 / 433 / 444 / 440 Mpaths/s; leaf batches of 8 / 24 / 32 lose 1-2%).
#ifndef TPT_REFILL_LANES
#define TPT_REFILL_LANES 24
#endif
#ifndef TPT_LEAF_BATCH
#define TPT_LEAF_BATCH 16
#endif
// leaf batches also start once fewer than this many lanes have inner nodes left
// (0: only at kLeafBatch parked leaves or no inner work).  Room 1024x768x256,
// depth 8: 0 / 16 / 24 give 612-619 / 618 / 622-624 Mpaths/s (extend alone 2.108
// / 2.060 / 2.067 ms per launch); testing a lane's second leaf in the same batch
// (instead of leaving the lane blocked until the next) lost 3%, refills at 16 or
// 32 idle lanes 2-3%, and leaf batches of 12 or 20 (with an inner minimum of 24,
// 28 or 32) stayed within 1%.  Spreading a batch's triangle tests over the wave
// (job slots by a prefix of the leaf sizes, owner lanes found by binary search
// over shuffles, a segmented (t, gid) min reduction; same images) raised the
// active lanes per VALU instruction from 33.7 to 40.4 but issued 5% more VALU
// instructions: leaves hold 1.86 triangles on average, so the loop below runs
// about two passes per batch and a batch is ~28% of the kernel's instructions,
// while the inner-node iterations (45.6 lanes busy on average) are ~65%.
// Re-checked with chunks of 512 entries: refill at 20 / 28 idle lanes, leaf
// batches of 20 and inner minimums of 20 / 28 stayed within 1% of these
// defaults (617-625 against 621-623 Mpaths/s), leaf batches of 12 lost 1.5%.
#ifndef TPT_LEAF_INNER_MIN
#define TPT_LEAF_INNER_MIN 24
#endif
static constexpr int kRefillLanes = TPT_REFILL_LANES;
static constexpr int kLeafBatch = TPT_LEAF_BATCH;
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

// Ray index sources of the refilling traversal loop.  take(k) returns the first
// of k consecutive ray indices for the wave's idle lanes (wave-uniform).
struct FetchGlobal { // one counter shared by every wave of the launch
    uint32_t *ctr;
    __device__ __forceinline__ uint32_t take(uint32_t k, int lane, int leader) {
        uint32_t base = 0;
        if (lane == leader) base = atomicAdd(ctr, k);
        return __shfl(base, leader);
    }
};
struct FetchLocal { // the wave's own range of the queue: no atomics, rays stay neighbours
    uint32_t cursor;
    __device__ __forceinline__ uint32_t take(uint32_t k, int, int) {
        uint32_t b = cursor;
        cursor += k;
        return b;
    }
};

// Loads ray i of `src` with its starting closest hit; true when the ray needs no
// traversal (an any-hit query already occluded, or no BVH).
template <bool ANY, class Src>
__device__ __forceinline__ bool ray_start(const DScene &S, const Src &src, uint32_t i, V3 &o, V3 &d, AlphaCtx &ac,
                                          HitRec &best) {
    if constexpr (Src::kAnalyticDone) {
        static_assert(!ANY, "precomputed analytic hits are closest hits");
        src.load_hit(S, i, o, d, ac, best);
        return S.root == -1;
    } else {
        float tmax;
        src.load(i, o, d, tmax, ac);
        best = HitRec{tmax, -1, 0.f, 0.f, 0};
        return analytic_tests<ANY>(S, o, d, ac, best) || S.root == -1;
    }
}

// Rays [.., n) from `fetch`; `cache`: LDS node-cache size (staged by the caller).
template <bool ANY, bool COUNT, int NL, bool POSTPONE, bool LDS, class Src, class Fetch>
__device__ __forceinline__ void persistent_trace(const DScene &S, const Src &src, uint32_t n, Fetch &fetch,
                                                 uint32_t &nodes, uint32_t &tris, int cache) {
    int32_t leaf = kNoNode; // parked leaf of this lane
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
            uint32_t base = fetch.take((uint32_t)nneed, lane, leader);
            if ((uint64_t)base + (uint64_t)nneed >= (uint64_t)n) exhausted = true;
            if ((need >> lane) & 1ull) {
                uint32_t idx = base + (uint32_t)__popcll(need & lower);
                if (idx < n) {
                    if (ray_start<ANY>(S, src, idx, o, d, ac, best)) {
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
                if (lm != 0 && (__popcll(lm) >= kLeafBatch || im == 0 ||
                                (TPT_LEAF_INNER_MIN > 0 && __popcll(im) < TPT_LEAF_INNER_MIN))) {
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
    uint32_t n = P.counters[qsel * kCntLine];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        P.counters[nsel_next * kCntLine] = 0u; // next ray queue
        P.counters[kCntShadow] = 0u;
        P.counters[kCntShFetch] = 0u;
        if (COUNT) atomicAdd(reinterpret_cast<unsigned long long *>(P.counters + kCntTraffic + 4), (unsigned long long)n);
    }
    uint32_t nodes = 0, tris = 0;
    SrcQueue src{queue, P.org, P.dir, P.key, P.depth, P.hit};
    const int cache = NL == kNodeBvh4Q ? S.lds_nodes : 0;
    stage_top_nodes(S, cache);
    FetchGlobal f{P.counters + kCntFetch};
    persistent_trace<false, COUNT, NL, POST, LDS>(S, src, n, f, nodes, tris, cache);
    if (COUNT) {
        wave_count(nodes, P.counters + kCntTraffic);
        wave_count(tris, P.counters + kCntTraffic + 2);
    }
}

// Entries per wave of a chunked launch over n rays: `chunk`, reduced in steps
// of one wave's width so that the rays still spread over at least `waves`
// wavefronts (late bounces and small renders have few rays).
__device__ __forceinline__ uint32_t chunk_of(uint32_t n, int chunk, int waves) {
    uint32_t c = (n + (uint32_t)waves - 1) / (uint32_t)waves;
    c = (c + 63u) / 64u * 64u;
    return max(64u, min((uint32_t)chunk, c));
}

// Chunked traversal: wave w owns queue entries [w * ck, (w + 1) * ck) and
// refills its idle lanes from them (FetchLocal), so lanes stay busy while the
// wave's rays stay neighbours in the queue; no shared counter.
// Occupancy of the chunked traversal kernels: 8 waves per SIMD (at most 64
// VGPRs; the default quantized kernel needs 66 otherwise and gets 7 waves).
// Room 1024x768x256: 655 -> 668 Mpaths/s, the extend kernel alone 4.85 -> 4.98
// Grays/s.  Occupancy-tuning builds may define it otherwise.
#ifndef TPT_CHUNK_ATTR
#define TPT_CHUNK_ATTR __attribute__((amdgpu_waves_per_eu(8)))
#endif

template <bool COUNT, int NL, bool POST, bool LDS>
__global__ TPT_CHUNK_ATTR void k_extend_chunk(DScene S, DPaths P, const int32_t *queue, int qsel, int nsel_next, int chunk,
                               int waves, int cache) {
    stage_top_nodes(S, cache);
    uint32_t n = P.counters[qsel * kCntLine];
    const uint32_t ck = chunk_of(n, chunk, waves);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        P.counters[nsel_next * kCntLine] = 0u; // next ray queue
        P.counters[kCntShadow] = 0u;
        if (COUNT) atomicAdd(reinterpret_cast<unsigned long long *>(P.counters + kCntTraffic + 4), (unsigned long long)n);
    }
    uint64_t begin = (uint64_t)((blockIdx.x * blockDim.x + threadIdx.x) >> 6) * (uint64_t)ck;
    if (begin >= n) return; // wave-uniform
    uint32_t end = (uint32_t)min(begin + (uint64_t)ck, (uint64_t)n);
    uint32_t nodes = 0, tris = 0;
    SrcQueue src{queue, P.org, P.dir, P.key, P.depth, P.hit};
    FetchLocal f{(uint32_t)begin};
    persistent_trace<false, COUNT, NL, POST, LDS>(S, src, end, f, nodes, tris, cache);
    if (COUNT) {
        wave_count(nodes, P.counters + kCntTraffic);
        wave_count(tris, P.counters + kCntTraffic + 2);
    }
}

template <bool COUNT, int NL, bool POST, bool LDS>
__global__ TPT_CHUNK_ATTR void k_shadow_chunk(DScene S, DPaths P, int chunk, int waves, int cache) {
    stage_top_nodes(S, cache);
    uint32_t n = P.counters[kCntShadow];
    if (COUNT && blockIdx.x == 0 && threadIdx.x == 0)
        atomicAdd(reinterpret_cast<unsigned long long *>(P.counters + kCntTraffic + 6), (unsigned long long)n);
    const uint32_t ck = chunk_of(n, chunk, waves);
    uint64_t begin = (uint64_t)((blockIdx.x * blockDim.x + threadIdx.x) >> 6) * (uint64_t)ck;
    if (begin >= n) return;
    uint32_t end = (uint32_t)min(begin + (uint64_t)ck, (uint64_t)n);
    uint32_t nodes = 0, tris = 0;
    SrcShadow src{P.shadow_org, P.shadow_dir, P.shadow_val, P.shadow_slot, P.color};
    FetchLocal f{(uint32_t)begin};
    persistent_trace<true, false, NL, POST, LDS>(S, src, end, f, nodes, tris, cache);
}

template <bool ANY, bool COUNT, int NL, bool POST, bool LDS>
__global__ TPT_CHUNK_ATTR void k_trace_chunk(DScene S, SrcRays src, uint32_t n, uint32_t *counters, int chunk, int waves,
                              int cache) {
    stage_top_nodes(S, cache);
    const uint32_t ck = chunk_of(n, chunk, waves);
    uint64_t begin = (uint64_t)((blockIdx.x * blockDim.x + threadIdx.x) >> 6) * (uint64_t)ck;
    if (begin >= n) return;
    uint32_t end = (uint32_t)min(begin + (uint64_t)ck, (uint64_t)n);
    uint32_t nodes = 0, tris = 0;
    FetchLocal f{(uint32_t)begin};
    persistent_trace<ANY, COUNT, NL, POST, LDS>(S, src, end, f, nodes, tris, cache);
    if (COUNT) {
        wave_count(nodes, counters + 0);
        wave_count(tris, counters + 2);
    }
}

template <bool COUNT, int NL, bool POST, bool LDS>
__global__ TPT_TRAVERSAL_ATTR void k_shadow_pt(DScene S, DPaths P) {
    uint32_t n = P.counters[kCntShadow];
    if (COUNT && blockIdx.x == 0 && threadIdx.x == 0)
        atomicAdd(reinterpret_cast<unsigned long long *>(P.counters + kCntTraffic + 6), (unsigned long long)n);
    uint32_t nodes = 0, tris = 0;
    SrcShadow src{P.shadow_org, P.shadow_dir, P.shadow_val, P.shadow_slot, P.color};
    const int cache = NL == kNodeBvh4Q ? S.lds_nodes : 0;
    stage_top_nodes(S, cache);
    FetchGlobal f{P.counters + kCntShFetch};
    persistent_trace<true, false, NL, POST, LDS>(S, src, n, f, nodes, tris, cache);
}

template <bool ANY, bool COUNT, int NL, bool POST, bool LDS>
__global__ TPT_TRAVERSAL_ATTR void k_trace_pt(DScene S, SrcRays src, uint32_t n, uint32_t *counters) {
    uint32_t nodes = 0, tris = 0;
    const int cache = NL == kNodeBvh4Q ? S.lds_nodes : 0;
    stage_top_nodes(S, cache);
    FetchGlobal f{counters + 8};
    persistent_trace<ANY, COUNT, NL, POST, LDS>(S, src, n, f, nodes, tris, cache);
    if (COUNT) {
        wave_count(nodes, counters + 0);
        wave_count(tris, counters + 2);
    }
}

// ---------------------------------------------------------------- generate
// Camera ray and erand48 stream of sample `s` of pixel `pix`, written to slot j.
// Writes a path ray with its closest analytic hit in the spare words (org.w = t,
// dir.w = analytic index * 2 + into, or -1), tested with the alpha context the
// extend kernels use for it (the path's key, the ray's depth).
__device__ __forceinline__ void put_ray(const DScene &S, const DPaths &P, int slot, V3 o, V3 d, uint64_t key,
                                        int depth) {
    HitRec a{kINF, -1, 0.f, 0.f, 0};
    if (S.num_analytic > 0) analytic_tests<false>(S, o, d, AlphaCtx{key, depth, true}, a);
    const int code = a.loc < 0 ? ((-a.loc - 1) << 1) | (a.u != 0.f ? 1 : 0) : -1;
    P.org[slot] = make_float4(o.x, o.y, o.z, a.t);
    P.dir[slot] = make_float4(d.x, d.y, d.z, __int_as_float(code));
}

// Camera ray of sample s of pixel pix with the path's stream seed `key` and its
// generator after the ray's draws.
__device__ __forceinline__ void camera_ray(const DScene &S, const DRender &R, int pix, int s, uint64_t base_mixed,
                                           V3 &org, V3 &dir, uint64_t &key, Rng &rng) {
    int W = S.cam.width;
    int x = pix % W, y = pix / W;
    key = stream_seed(base_mixed, (uint64_t)pix, (uint64_t)s);
    rng = Rng{rng_state_from_seed(key)};
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
    dir = sub(add(add(scl(a, mk(c.right)), scl(b, mk(c.up))), mk(c.direct)), offset);
    dir = nrm(dir);
    org = add(mk(c.center), offset);
}

__device__ __forceinline__ void gen_path(const DScene &S, const DRender &R, const DPaths &P, int j, int pix, int s,
                                         uint64_t base_mixed) {
    V3 org, dir;
    uint64_t key;
    Rng rng;
    camera_ray(S, R, pix, s, base_mixed, org, dir, key, rng);
    put_ray(S, P, j, org, dir, key, 0);
    P.decay[j] = make_float4(1.f, 1.f, 1.f, 0.f);
    P.color[j] = make_float4(0.f, 0.f, 0.f, 0.f);
    P.rng[j] = rng.x;
    P.key[j] = key;
    P.depth[j] = 0;
    P.queue_a[j] = j;
}

__global__ void k_generate(DScene S, DRender R, DPaths P, int wave_base, int n_paths, uint64_t base_mixed) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j == 0) {
        P.counters[kCntQueue] = (uint32_t)n_paths;
        P.counters[kCntFetch] = 0u;
    }
    if (j >= n_paths) return;
    gen_path(S, R, P, j, j % R.pixels, wave_base + j / R.pixels, base_mixed);
}

// ---------------------------------------------------------------- extend
// One ray per lane: a wavefront takes 64 consecutive queue entries and keeps them
// until all 64 are done.  Neighbouring queue entries are mostly neighbouring
// pixels, so the lanes of a wave tend to fetch the same nodes and triangles and
// their loads coalesce.  `cache`: LDS node-cache size (stage_top_nodes).
template <bool COUNT, int NL, bool LDS>
__global__ void k_extend(DScene S, DPaths P, const int32_t *queue, int qsel, int nsel_next, int cache) {
    stage_top_nodes(S, cache);
    uint32_t n = P.counters[qsel * kCntLine];
    int tid = blockIdx.x * blockDim.x + threadIdx.x;
    if (tid == 0) {
        P.counters[nsel_next * kCntLine] = 0u; // next ray queue
        P.counters[kCntShadow] = 0u;
        if (COUNT) atomicAdd(reinterpret_cast<unsigned long long *>(P.counters + kCntTraffic + 4), (unsigned long long)n);
    }
    uint32_t nodes = 0, tris = 0;
    for (uint32_t i = (uint32_t)tid; i < n; i += gridDim.x * blockDim.x) {
        int slot = queue[i];
        float4 o4 = P.org[slot], d4 = P.dir[slot];
        AlphaCtx ac{P.key[slot], P.depth[slot], true};
        HitRec h = trace<false, COUNT, NL, LDS>(S, mk4(o4), mk4(d4), kINF, ac, nodes, tris, cache);
        P.hit[i] = make_float4(h.t, __int_as_float(h.loc), h.u, h.v);
    }
    if (COUNT) {
        wave_count(nodes, P.counters + kCntTraffic);
        wave_count(tris, P.counters + kCntTraffic + 2);
    }
}

// ---------------------------------------------------------------- shade
__device__ __forceinline__ void surface(const DScene &S, V3 o, V3 d, const HitRec &h, SurfHit &sh, int &mat) {
    int32_t loc = h.loc;
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
    // triangle (object.cpp:266-331), from its shading record (one cache line); the
    // material index comes with the location when it fits there, so its record is
    // read beside the shading record instead of after it
    const float4 *sr = S.shade_rec + (size_t)kShadeRec * loc_tri(S, loc);
    const float4 r0 = sr[0], r1 = sr[1], r2 = sr[2], r3 = sr[3], r4 = sr[4], r5 = sr[5];
    mat = S.loc_bits ? (int)((uint32_t)loc >> S.loc_bits) : __float_as_int(r0.w);
    const int flags = __float_as_int(r1.w);
    float u = h.u, v = h.v;
    V3 ab = mk4(r0), ac = mk4(r1);
    const V2 ta{r4.y, r4.z}, tb{r4.w, r5.x}, tcv{r5.y, r5.z};
    V2 uv;
    if (flags & 2) {
        float w0 = 1.f - u - v;
        uv = {(w0 * ta.x + u * tb.x) + v * tcv.x, (w0 * ta.y + u * tb.y) + v * tcv.y};
    } else {
        uv = {u, v};
    }
    V3 n;
    if (flags & 1) {
        float w0 = 1.f - u - v;
        V3 na{r2.x, r2.y, r2.z}, nb{r2.w, r3.x, r3.y}, nc{r3.z, r3.w, r4.x};
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
        if (flags & 2) {
            dab = {tb.x - ta.x, tb.y - ta.y};
            dac = {tcv.x - ta.x, tcv.y - ta.y};
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

// One path vertex (the loop body of cpu/path_tracer.cpp:102-168) for the ray at
// queue position i (slot `slot`): adds emission / background to the path's
// radiance, samples the next direction and, for a non-specular lobe with delta
// lights, prepares one shadow ray.  `alive`: the path continues.
struct ShadeOut {
    bool alive, emit_shadow;
    V3 s_org, s_dir, s_val;
    float s_tmax;
    int cache_kind; // 0: none, 1: terminated into the radiance cache, 2: training record
};

// Roughness and albedo of a material's first lobe (radiance-cache inputs).
__device__ __forceinline__ void lobe_summary(const DScene &S, const DMaterial &m, V2 uv, float &rough, V3 &alb) {
    const DLobe &b = m.lobe[0];
    rough = b.type == 0 ? 1.f : (b.type == 1 || b.type == 2) ? 0.f : sqrtf(b.alpha2);
    alb = color_at(S, b.color, uv);
}

// One training path in `stride` per frame, by a hash of (frame, slot).
__device__ __forceinline__ bool cache_trains(const DCache &C, int slot) {
    return splitmix64(((uint64_t)C.frame << 32) ^ (uint64_t)(uint32_t)slot) % (uint64_t)C.train_stride == 0;
}

// The state of the path being shaded, read and written through `Path`: PoolPath
// (the wavefront kernels: the path pool, the hit record at queue position i) or
// RegPath (k_megapath: registers).  Returns the shadow ray to trace, if any.
template <class Path>
__device__ __forceinline__ ShadeOut shade_path(const DScene &S, const DRender &R, Path &st, const DCache *C) {
    const int slot = st.slot;
    // cpu/path_tracer.cpp:104; cached frames may sample the lights more often
    const float LIGHT_PROB = C ? C->light_prob : 1.f / 5.f;
    ShadeOut out_{false, false, V3{0, 0, 0}, V3{0, 0, 0}, V3{0, 0, 0}, kINF, 0};
    bool &alive = out_.alive, &emit_shadow = out_.emit_shadow;
    V3 &s_org = out_.s_org, &s_dir = out_.s_dir, &s_val = out_.s_val;
    float &s_tmax = out_.s_tmax;
    HitRec h = st.hit();
    V3 o = st.org(), d = st.dir();
    V3 decay = st.decay();
    V3 color = st.color();
    int depth = st.depth();
    Rng rng{st.rng()};
    if (!(h.t < kINF)) {
        color = add(color, cmul(decay, background(S, d)));
    } else {
        SurfHit sh;
        int mi;
        surface(S, o, d, h, sh, mi);
        const DMaterial &m = S.materials[mi];
        color = add(color, cmul(decay, color_at(S, m.emission, sh.uv)));
        bool into_cache = false;
        // The cache vertex: the first vertex of depth >= C->depth whose first lobe is
        // rough enough (smooth mirrors and glass send the path on, Mueller et al.
        // 2021).  A path ends there and queries the cache; a training path records
        // the vertex and goes on, and queries the cache `suffix` vertices later.
        if (C && depth >= C->depth) {
            float rough;
            V3 alb;
            lobe_summary(S, m, sh.uv, rough, alb);
            const bool eligible = rough >= C->min_rough;
            const bool trains = cache_trains(*C, slot);
            // depth of the training path's recorded vertex (none yet at the first candidate depth)
            const int rec = trains && depth > C->depth ? (int)C->rec_T[slot].w : -1;
            const bool record = trains && rec < 0 && eligible;
            const bool query = (!trains && eligible) || (trains && rec >= 0 && C->suffix > 0 && depth == rec + C->suffix);
            if (record || query) {
                V3 hp = at(o, d, h.t);
                V3 p01{(hp.x - C->lo[0]) * C->inv_ext[0], (hp.y - C->lo[1]) * C->inv_ext[1],
                       (hp.z - C->lo[2]) * C->inv_ext[2]};
                nrc_encode((record ? C->feat_t : C->feat) + (size_t)slot * kNrcInputs, p01, neg(d), sh.shade_normal,
                           rough, alb);
            }
            // with the albedo factor, the throughput into the vertex includes the albedo:
            // the recorded target becomes radiance / albedo and a query's prediction
            // is multiplied by it
            const bool factor = C->albedo_factor == 2 || (C->albedo_factor == 1 && m.lobe[0].color.image >= 0);
            const V3 tf = factor ? cmul(decay, alb) : decay;
            if (record) {
                out_.cache_kind = 2;
                C->rec_T[slot] = make_float4(tf.x, tf.y, tf.z, (float)depth);
                C->rec_C[slot] = make_float4(color.x, color.y, color.z, 0.f);
            } else if (query) {
                out_.cache_kind = 1;
                into_cache = true; // P.decay[slot]: the throughput the prediction is applied with
                if (factor) st.set_decay(tf);
            } else if (trains && depth == C->depth) {
                C->rec_T[slot] = make_float4(0.f, 0.f, 0.f, -1.f); // nothing recorded yet
            }
        }
        if (depth < R.max_depth && !into_cache) {
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
                st.next_ray(S, hp, out, depth + 1);
                st.set_decay(decay);
                st.set_depth(depth + 1);
            }
        }
    }
    st.set_color(color);
    st.set_rng(rng.x);
    return out_;
}

// Path state in the path pool; the hit record of the path at queue position i.
struct PoolPath {
    const DPaths &P;
    uint32_t i;
    int slot;
    __device__ __forceinline__ HitRec hit() const {
        float4 h4 = P.hit[i];
        return HitRec{h4.x, -1, h4.z, h4.w, __float_as_int(h4.y)}; // the path pipeline's hit records carry the local id
    }
    __device__ __forceinline__ V3 org() const { return mk4(P.org[slot]); }
    __device__ __forceinline__ V3 dir() const { return mk4(P.dir[slot]); }
    __device__ __forceinline__ V3 decay() const { return mk4(P.decay[slot]); }
    __device__ __forceinline__ V3 color() const { return mk4(P.color[slot]); }
    __device__ __forceinline__ int depth() const { return P.depth[slot]; }
    __device__ __forceinline__ uint64_t rng() const { return P.rng[slot]; }
    __device__ __forceinline__ void next_ray(const DScene &S, V3 o, V3 d, int depth) const {
        put_ray(S, P, slot, o, d, P.key[slot], depth);
    }
    __device__ __forceinline__ void set_decay(V3 v) const { P.decay[slot] = make_float4(v.x, v.y, v.z, 0.f); }
    __device__ __forceinline__ void set_depth(int v) const { P.depth[slot] = v; }
    __device__ __forceinline__ void set_color(V3 c) const { P.color[slot] = make_float4(c.x, c.y, c.z, 0.f); }
    __device__ __forceinline__ void set_rng(uint64_t x) const { P.rng[slot] = x; }
};

__device__ __forceinline__ ShadeOut shade_entry(const DScene &S, const DRender &R, const DPaths &P, uint32_t i,
                                                int slot, const DCache *C = nullptr) {
    PoolPath pp{P, i, slot};
    return shade_path(S, R, pp, C);
}

// Extra attributes of the shade kernels (occupancy-tuning builds define it, e.g.
// as __attribute__((amdgpu_waves_per_eu(6)))).
#ifndef TPT_SHADE_ATTR
#define TPT_SHADE_ATTR
#endif

// CACHE: the radiance-cache hooks are on; queries and training records are
// appended to C's lists (at most C.train_cap training records).
// Entries are shaded in queue order.  Sorting each workgroup's entries by
// material class first (class bits in the hit record, an LDS counting sort; same
// images) was slower on room 1024x768x256: the next queue inherits the sorted
// order, so extend's lanes lose their ray coherence (2.05 -> 2.20 ms per
// launch), and the sort's LDS and barriers cost shade more than the lane
// divergence it removes (209 -> 219 ms per render alone; three streams 322 ->
// 380 ms, the shade workgroups' LDS now competing with extend's stacks).
template <bool CACHE>
__global__ TPT_SHADE_ATTR void k_shade(DScene S, DRender R, DPaths P, const int32_t *queue, int32_t *next, int qsel, int nsel,
                        DCache C) {
    uint32_t n = P.counters[qsel * kCntLine];
    if (blockIdx.x == 0 && threadIdx.x == 0) P.counters[kCntFetch] = 0u; // refill counter of the next extend
    for (uint32_t base = blockIdx.x * blockDim.x; base < n; base += gridDim.x * blockDim.x) {
        uint32_t i = base + threadIdx.x;
        bool valid = i < n;
        int slot = valid ? queue[i] : 0;
        ShadeOut r{false, false, V3{0, 0, 0}, V3{0, 0, 0}, V3{0, 0, 0}, kINF, 0};
        if (valid) r = shade_entry(S, R, P, i, slot, CACHE ? &C : nullptr);
        if (CACHE) {
            int qp = wave_append(r.cache_kind == 1, C.counters + 0);
            if (r.cache_kind == 1) C.q_slot[qp] = slot;
            int tp = wave_append(r.cache_kind == 2, C.counters + 1);
            if (r.cache_kind == 2 && tp < C.train_cap) C.t_slot[tp] = slot;
        }
        int pos = wave_append(r.alive, P.counters + nsel * kCntLine);
        if (r.alive) next[pos] = slot;
        int spos = wave_append(r.emit_shadow, P.counters + kCntShadow);
        if (r.emit_shadow) {
            P.shadow_slot[spos] = slot;
            P.shadow_org[spos] = make_float4(r.s_org.x, r.s_org.y, r.s_org.z, r.s_tmax);
            P.shadow_dir[spos] = make_float4(r.s_dir.x, r.s_dir.y, r.s_dir.z, 0.f);
            P.shadow_val[spos] = make_float4(r.s_val.x, r.s_val.y, r.s_val.z, 0.f);
        }
    }
}

// ---------------------------------------------------------------- whole paths
// Path state in registers (k_megapath).
struct RegPath {
    int slot;
    HitRec h;
    V3 o, d, dec, col;
    int dep;
    uint64_t rs, key;
    __device__ __forceinline__ HitRec hit() const { return h; }
    __device__ __forceinline__ V3 org() const { return o; }
    __device__ __forceinline__ V3 dir() const { return d; }
    __device__ __forceinline__ V3 decay() const { return dec; }
    __device__ __forceinline__ V3 color() const { return col; }
    __device__ __forceinline__ int depth() const { return dep; }
    __device__ __forceinline__ uint64_t rng() const { return rs; }
    __device__ __forceinline__ void next_ray(const DScene &, V3 no, V3 nd, int) {
        o = no;
        d = nd;
    }
    __device__ __forceinline__ void set_decay(V3 v) { dec = v; }
    __device__ __forceinline__ void set_depth(int v) { dep = v; }
    __device__ __forceinline__ void set_color(V3 c) { col = c; }
    __device__ __forceinline__ void set_rng(uint64_t x) { rs = x; }
};

// Scenes of analytic primitives only (no BVH): each thread traces path j of the
// wave from its camera ray to its end with the state in registers (the shape of
// the reference's CUDA kernel, cuda/path_tracer_kernel.cu:80), instead of passing it
// through the pool between the wavefront kernels at every bounce.  A ray's
// closest hit is the analytic list's (put_ray's test, same alpha context), a
// shadow ray's occlusion the list's any-hit test (the shadow kernels' ray_start),
// and the radiance is added in the wavefront's order (the vertex's emission and
// BxDF sample in shade_path, then its shadow ray's contribution), so each path's
// colour equals the wavefront kernels' bit for bit.  Writes it to P.color[j] for
// k_accumulate.  REFILL: a lane whose path has ended starts its next path (j +
// the grid's thread count) instead of idling until the wave's longest path ends;
// the idle lanes start at the top of the loop, once refill_lanes of them are idle
// (a wave-uniform decision).  Two-spheres box 512x384x64 (depth 10, RR 5; same
// images): one thread per path 4.65 ms per render, refill at 1 / 4 / 16 idle
// lanes 4.15 / 4.25-4.30 / 4.38 ms; a lane restarting inside the bounce's
// path-ended branch instead took 5.0-8.3 ms.
__device__ __forceinline__ void mega_start(const DScene &S, const DRender &R, int j, int wave_base,
                                           uint64_t base_mixed, RegPath &st) {
    Rng rng;
    camera_ray(S, R, j % R.pixels, wave_base + j / R.pixels, base_mixed, st.o, st.d, st.key, rng);
    st.slot = j;
    st.rs = rng.x;
    st.dec = V3{1.f, 1.f, 1.f};
    st.col = V3{0.f, 0.f, 0.f};
    st.dep = 0;
}
// One vertex of the path in `st` (closest hit, shading, shadow ray); false once the path has ended.
__device__ __forceinline__ bool mega_bounce(const DScene &S, const DRender &R, RegPath &st) {
    HitRec a{kINF, -1, 0.f, 0.f, 0};
    analytic_tests<false>(S, st.o, st.d, AlphaCtx{st.key, st.dep, true}, a);
    st.h = HitRec{a.t, -1, a.u, a.v, a.loc};
    const ShadeOut r = shade_path(S, R, st, nullptr);
    if (r.emit_shadow) {
        HitRec b{r.s_tmax, -1, 0.f, 0.f, 0};
        if (!analytic_tests<true>(S, r.s_org, r.s_dir, AlphaCtx{0, 0, false}, b))
            st.col = V3{st.col.x + r.s_val.x, st.col.y + r.s_val.y, st.col.z + r.s_val.z};
    }
    return r.alive;
}
template <bool REFILL>
__global__ TPT_SHADE_ATTR void k_megapath(DScene S, DRender R, DPaths P, int wave_base, int n_paths,
                                          uint64_t base_mixed, int refill_lanes) {
    const int stride = gridDim.x * blockDim.x;
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (REFILL) {
        // idle lanes start their next paths together, once at least refill_lanes
        // of the wave's lanes are idle or none is tracing
        RegPath st;
        bool live = false;
        while (true) {
            const uint64_t idle = __ballot(!live && j < n_paths);
            const uint64_t act = __ballot(live);
            if (idle == 0 && act == 0) break;
            if (__popcll(idle) >= refill_lanes || act == 0) {
                if (!live && j < n_paths) {
                    mega_start(S, R, j, wave_base, base_mixed, st);
                    live = true;
                }
            }
            if (live && !mega_bounce(S, R, st)) {
                P.color[j] = make_float4(st.col.x, st.col.y, st.col.z, 0.f);
                live = false;
                j += stride;
            }
        }
    } else {
        for (; j < n_paths; j += stride) {
            RegPath st;
            mega_start(S, R, j, wave_base, base_mixed, st);
            while (mega_bounce(S, R, st)) {
            }
            P.color[j] = make_float4(st.col.x, st.col.y, st.col.z, 0.f);
        }
    }
}

// ---------------------------------------------------------------- shadow
template <bool COUNT, int NL, bool LDS>
__global__ void k_shadow(DScene S, DPaths P, int cache) {
    stage_top_nodes(S, cache);
    uint32_t n = P.counters[kCntShadow];
    if (COUNT && blockIdx.x == 0 && threadIdx.x == 0)
        atomicAdd(reinterpret_cast<unsigned long long *>(P.counters + kCntTraffic + 6), (unsigned long long)n);
    uint32_t nodes = 0, tris = 0;
    AlphaCtx ac{0, 0, false};
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        float4 o4 = P.shadow_org[i], d4 = P.shadow_dir[i];
        HitRec h = trace<true, false, NL, LDS>(S, mk4(o4), mk4(d4), o4.w, ac, nodes, tris, cache);
        if (h.gid < 0) {
            int slot = P.shadow_slot[i];
            float4 c = P.color[slot], v = P.shadow_val[i];
            P.color[slot] = make_float4(c.x + v.x, c.y + v.y, c.z + v.z, 0.f);
        }
    }
}

// ---------------------------------------------------------------- accumulate / finalize
// Adds `count` consecutive samples (s_first, s_first + 1, ...) of one pixel to its
// running sums at accumulator index q, in sample order; `color_slot(j)` is the
// path slot of sample s_first + j.  A sub-pixel closes after nsub samples
// (cpu/path_tracer.cpp:80-96): its mean is clipped (LDR) and added to the pixel.
template <class SlotOf>
__device__ __forceinline__ void accum_run(const DRender &R, const DPaths &P, size_t q, int s_first, int count,
                                          SlotOf color_slot, bool first_wave) {
    float sf[3], pf[3];
    double sd[3], pd[3];
    for (int c = 0; c < 3; c++) {
        sf[c] = first_wave ? 0.f : P.sub_f[3 * q + c];
        pf[c] = first_wave ? 0.f : P.pix_f[3 * q + c];
        sd[c] = (first_wave || !R.accum_fp64) ? 0.0 : P.sub_d[3 * q + c];
        pd[c] = (first_wave || !R.accum_fp64) ? 0.0 : P.pix_d[3 * q + c];
    }
    for (int j = 0; j < count; j++) {
        int s = s_first + j;
        float4 L = P.color[color_slot(j)];
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
        P.sub_f[3 * q + c] = sf[c];
        P.pix_f[3 * q + c] = pf[c];
        if (R.accum_fp64) {
            P.sub_d[3 * q + c] = sd[c];
            P.pix_d[3 * q + c] = pd[c];
        }
    }
}

__global__ void k_accumulate(DRender R, DPaths P, int wave_base, int k_active, int first_wave) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= R.pixels) return;
    accum_run(R, P, (size_t)p, wave_base, k_active, [&](int j) { return (size_t)j * R.pixels + p; }, first_wave != 0);
}

// Output value of channel c at accumulator index q.
__device__ __forceinline__ float final_value(const DRender &R, const DPaths &P, size_t q, int c) {
    int bb = R.block * R.block;
    if (R.output == 1) {
        if (R.accum_fp64) return (float)(P.pix_d[3 * q + c] / ((double)R.nsub * (double)bb));
        return P.pix_f[3 * q + c] / (float)(R.nsub * bb);
    }
    return clip1(P.pix_f[3 * q + c] / (float)bb);
}

__global__ void k_finalize(DRender R, DPaths P, float *out) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= R.pixels) return;
    for (int c = 0; c < 3; c++) out[3 * p + c] = final_value(R, P, (size_t)p, c);
}

// ---------------------------------------------------------------- ray queries (parity / bench)
template <bool COUNT, int NL, bool LDS>
__global__ void k_trace_closest(DScene S, const float4 *rays, uint64_t n, int32_t *prim, float *tout,
                                uint32_t *counters, int cache) {
    stage_top_nodes(S, cache);
    uint32_t nodes = 0, tris = 0;
    AlphaCtx ac{0, 0, false};
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        float4 a = rays[2 * i], b = rays[2 * i + 1];
        float tmax = b.w < kINF ? b.w : kINF;
        HitRec h = trace<false, COUNT, NL, LDS>(S, mk4(a), mk4(b), tmax, ac, nodes, tris, cache);
        bool hit = h.gid >= 0;
        prim[i] = hit ? h.gid : -1;
        tout[i] = hit ? h.t : kINF;
    }
    if (COUNT) {
        wave_count(nodes, counters + 0);
        wave_count(tris, counters + 2);
    }
}

template <int NL, bool LDS>
__global__ void k_trace_occluded(DScene S, const float4 *rays, uint64_t n, uint8_t *occ, int cache) {
    stage_top_nodes(S, cache);
    uint32_t nodes = 0, tris = 0;
    AlphaCtx ac{0, 0, false};
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        float4 a = rays[2 * i], b = rays[2 * i + 1];
        float tmax = b.w < kINF ? b.w : kINF;
        HitRec h = trace<true, false, NL, LDS>(S, mk4(a), mk4(b), tmax, ac, nodes, tris, cache);
        occ[i] = h.gid >= 0 ? 1 : 0;
    }
}

// Sets n words to v: clears counters with a kernel (no memset node when a caller
// captures the work in a graph).
__global__ void k_fill_u32(uint32_t *p, int n, uint32_t v) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// ---------------------------------------------------------------- radiance cache (progressive frames)
__global__ void k_cache_reset(DCache C) {
    if (threadIdx.x < 2) C.counters[threadIdx.x] = 0u;
}

// color += decay * max(prediction, 0) for every path that terminated into the cache
__global__ void k_cache_apply(DPaths P, DCache C, const float4 *pred, int n_max) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int n = min((int)C.counters[0], n_max);
    if (i >= n) return;
    int s = C.q_slot[i];
    float4 p = pred[i], d = P.decay[s], c = P.color[s];
    P.color[s] = make_float4(c.x + d.x * fmaxf(p.x, 0.f), c.y + d.y * fmaxf(p.y, 0.f), c.z + d.z * fmaxf(p.z, 0.f), 0.f);
}

// Training targets: the radiance each recorded vertex scattered towards the
// camera, (final radiance - radiance after the vertex's emission) / throughput.
__device__ __forceinline__ float cache_target(float fin, float at, float T) {
    if (!(T > 1e-6f)) return 0.f;
    return fminf(fmaxf((fin - at) / T, 0.f), 1e4f);
}
__global__ void k_cache_targets(DPaths P, DCache C, float4 *target) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int n = min((int)C.counters[1], C.train_cap);
    if (i >= n) return;
    int s = C.t_slot[i];
    float4 f = P.color[s], a = C.rec_C[s], T = C.rec_T[s];
    target[i] = make_float4(cache_target(f.x, a.x, T.x), cache_target(f.y, a.y, T.y), cache_target(f.z, a.z, T.z), 0.f);
}

// `frames` one-sample-per-pixel frames (frame g's pixel p in slot g * pixels + p)
// added to the running per-pixel sums one after the other, in frame order.
__global__ void k_frame_accumulate(DPaths P, int pixels, int fp64, int first, int frames) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= pixels) return;
    if (fp64) {
        double s[3];
        for (int c = 0; c < 3; c++) s[c] = first ? 0.0 : P.pix_d[3 * p + c];
        for (int g = 0; g < frames; g++) {
            const float4 L = P.color[(size_t)g * pixels + p];
            s[0] += (double)L.x;
            s[1] += (double)L.y;
            s[2] += (double)L.z;
        }
        for (int c = 0; c < 3; c++) P.pix_d[3 * p + c] = s[c];
    } else {
        float s[3];
        for (int c = 0; c < 3; c++) s[c] = first ? 0.f : P.pix_f[3 * p + c];
        for (int g = 0; g < frames; g++) {
            const float4 L = P.color[(size_t)g * pixels + p];
            s[0] += L.x;
            s[1] += L.y;
            s[2] += L.z;
        }
        for (int c = 0; c < 3; c++) P.pix_f[3 * p + c] = s[c];
    }
}
__global__ void k_frame_finalize(DPaths P, int pixels, int fp64, int frames, float *out) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= pixels) return;
    for (int c = 0; c < 3; c++)
        out[3 * p + c] = fp64 ? (float)(P.pix_d[3 * p + c] / (double)frames) : P.pix_f[3 * p + c] / (float)frames;
}

// ---------------------------------------------------------------- batch mode
// Many independent scenes share one path pool: the kernels below are the batch
// counterparts of generate / extend / shade / shadow / accumulate, built on the
// same per-path device functions, so every scene's image equals its
// single-scene render bit for bit.
#define TPT_SEG(B, f, s) (B).seg[(f) * (B).seg_stride + (s)]

// Scene s of the batch.  The table stores each array as a byte offset from
// B.base (a kernel argument): a pointer read from memory would be generic and
// compile to flat loads, while base + offset stays a global pointer, so every
// material, texture and node access is a global load as in the single-scene
// kernels.
template <class T> __device__ __forceinline__ T *rebase(const uint8_t *base, T *off) {
    return (T *)(base + (uintptr_t)off);
}
// The batch scene and sampler tables are read with load_invariant (dev_scene.h).
__device__ __forceinline__ DRender render_at(const DBatch &B, int s) { return load_invariant(B.renders + s); }
__device__ __forceinline__ DScene scene_at(const DBatch &B, int s) {
    DScene S = load_invariant(B.scenes + s);
    S.nodes = rebase(B.base, S.nodes);
    S.tris_bvh = rebase(B.base, S.tris_bvh);
    S.tris = rebase(B.base, S.tris);
    S.tri_info = rebase(B.base, S.tri_info);
    S.shade_rec = rebase(B.base, S.shade_rec);
    S.normals = rebase(B.base, S.normals);
    S.texcoords = rebase(B.base, S.texcoords);
    S.gid_local = rebase(B.base, S.gid_local);
    S.analytic = rebase(B.base, S.analytic);
    S.materials = rebase(B.base, S.materials);
    S.images = rebase(B.base, S.images);
    S.texels = rebase(B.base, S.texels);
    S.delta = rebase(B.base, S.delta);
    S.lights = rebase(B.base, S.lights);
    S.env_marg = rebase(B.base, S.env_marg);
    S.env_cond = rebase(B.base, S.env_cond);
    S.env_gmarg = rebase(B.base, S.env_gmarg);
    S.env_gcond = rebase(B.base, S.env_gcond);
    S.light_shapes = rebase(B.base, S.light_shapes);
    return S;
}

// Largest s in [0, n) with off[s] <= v (off non-decreasing, off[0] <= v).
template <class T> __device__ __forceinline__ int find_seg(const T *off, int n, T v) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (off[mid] <= v) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__device__ __forceinline__ uint64_t batch_wave_paths(const DBatch &B) {
    uint64_t g0 = B.state[0], total = B.state[1];
    return total > g0 ? min((uint64_t)B.cap, total - g0) : 0;
}

// Ray-queue segments of the pipeline wave starting at g0: scene s owns slots
// [path_off[s] - g0, path_off[s + 1] - g0) clipped to the wave.
__global__ void k_batch_begin(DBatch B) {
    uint64_t g0 = B.state[0], nw = batch_wave_paths(B);
    for (int s = threadIdx.x; s < B.n; s += blockDim.x) {
        uint64_t lo = B.path_off[s] > g0 ? min(B.path_off[s] - g0, nw) : 0;
        uint64_t hi = B.path_off[s + 1] > g0 ? min(B.path_off[s + 1] - g0, nw) : 0;
        TPT_SEG(B, kSegQaBase, s) = (int32_t)lo;
        TPT_SEG(B, kSegQaCnt, s) = (int32_t)(hi - lo);
        TPT_SEG(B, kSegFetch, s) = 0;
    }
}

__global__ void k_generate_batch(DBatch B, DPaths P) {
    uint64_t j = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (j >= batch_wave_paths(B)) return;
    uint64_t g = B.state[0] + j;
    int s = find_seg(B.path_off, B.n, g);
    const DRender &R = B.renders[s];
    uint64_t l = g - B.path_off[s];
    gen_path(scene_at(B, s), R, P, (int)j, (int)(l % (uint64_t)R.pixels), (int)(l / (uint64_t)R.pixels),
             B.seed_mixed[s]);
    P.sid[j] = s;
}

// Persistent traversal over the per-scene segments: each wavefront starts at its
// own scene (spreading the wavefronts over the batch) and walks all scenes,
// draining one scene's segment before moving to the next, so the scene is
// uniform across the wavefront and its record is read through scalar loads.
__device__ __forceinline__ int batch_start_scene(const DBatch &B) {
    int wave = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
    int nwaves = (int)((gridDim.x * blockDim.x) >> 6);
    return (int)(((int64_t)wave * B.n) / nwaves);
}

template <int NL, bool LDS>
__global__ TPT_TRAVERSAL_ATTR void k_extend_batch(DBatch B, DPaths P, int qsel) {
    const int fb = qsel ? kSegQbBase : kSegQaBase;
    const int32_t *queue = qsel ? P.queue_b : P.queue_a;
    const int start = batch_start_scene(B);
    uint32_t nodes = 0, tris = 0;
    for (int k = 0; k < B.n; k++) {
        int s = start + k < B.n ? start + k : start + k - B.n;
        s = __builtin_amdgcn_readfirstlane(s);
        int cnt = TPT_SEG(B, fb + 1, s);
        if (cnt == 0) continue;
        int base = TPT_SEG(B, fb, s);
        SrcQueue src{queue + base, P.org, P.dir, P.key, P.depth, P.hit + base};
        FetchGlobal f{(uint32_t *)&TPT_SEG(B, kSegFetch, s)};
        persistent_trace<false, false, NL, true, LDS>(scene_at(B, s), src, (uint32_t)cnt, f, nodes, tris, 0);
    }
}

// Exclusive scan of get(0 .. n) by one wavefront (each lane walks a run of
// ceil(n / 64) entries): put(i, prefix) for every i; returns the total.  The
// per-bounce scans of the batch pipeline run as single wavefronts so that they
// start at once even while another pipeline wave's kernels occupy every CU.
template <class Get, class Put> __device__ __forceinline__ int wave_scan(int n, Get get, Put put) {
    const int lane = (int)__lane_id(), per = (n + 63) / 64;
    const int lo = min(n, lane * per), hi = min(n, lo + per);
    int local = 0;
    for (int i = lo; i < hi; i++) local += get(i);
    int incl = local;
    for (int off = 1; off < 64; off <<= 1) {
        int y = __shfl_up(incl, off);
        if (lane >= off) incl += y;
    }
    const int total = __shfl(incl, 63);
    int run = incl - local;
    for (int i = lo; i < hi; i++) {
        int v = get(i);
        put(i, run);
        run += v;
    }
    return total;
}

// Next-bounce segments: the survivors (and shadow rays) of scene s go to
// [scan_s, scan_s + count_s) where scan is the exclusive prefix sum of this
// bounce's per-scene ray counts (a path leaves at most one survivor and one
// shadow ray).  One wavefront.
__global__ void k_batch_seg(DBatch B, int qsel) {
    const int cc = (qsel ? kSegQbBase : kSegQaBase) + 1;
    const int nb = qsel ? kSegQaBase : kSegQbBase;
    wave_scan(B.n, [&](int s) { return TPT_SEG(B, cc, s); }, [&](int s, int excl) {
        TPT_SEG(B, nb, s) = excl;
        TPT_SEG(B, nb + 1, s) = 0;
        TPT_SEG(B, kSegShBase, s) = excl;
        TPT_SEG(B, kSegShCnt, s) = 0;
        TPT_SEG(B, kSegShFetch, s) = 0;
        TPT_SEG(B, kSegFetch, s) = 0;
    });
}

// Workgroups are at most 512 threads (the autotuner's range), which lets the
// compiler budget registers for that size instead of 1024.
__global__ __attribute__((amdgpu_flat_work_group_size(64, 512))) TPT_SHADE_ATTR void
k_shade_batch(DBatch B, DPaths P, int qsel) {
    const int cb = qsel ? kSegQbBase : kSegQaBase, nb = qsel ? kSegQaBase : kSegQbBase;
    const int32_t *queue = qsel ? P.queue_b : P.queue_a;
    int32_t *next = qsel ? P.queue_a : P.queue_b;
    const int32_t *cbase = B.seg + cb * B.seg_stride;
    for (uint32_t base = blockIdx.x * blockDim.x; base < (uint32_t)B.cap; base += gridDim.x * blockDim.x) {
        uint32_t i = base + threadIdx.x;
        int s = 0;
        bool valid = false;
        if (i < (uint32_t)B.cap) {
            s = find_seg(cbase, B.n, (int32_t)i);
            valid = (int32_t)i < cbase[s] + TPT_SEG(B, cb + 1, s);
        }
        int slot = valid ? queue[i] : 0;
        // one pass per distinct scene among the wavefront's entries (usually one)
        uint64_t todo = __ballot(valid);
        while (todo) {
            int lead = __ffsll((unsigned long long)todo) - 1;
            int s0 = __builtin_amdgcn_readlane(s, lead);
            bool mine = valid && s == s0;
            ShadeOut r{false, false, V3{0, 0, 0}, V3{0, 0, 0}, V3{0, 0, 0}, kINF, 0};
            if (mine) {
                // `mine` implies s == s0, and the compiler may then index the tables
                // with the per-lane s; readfirstlane keeps the index in an SGPR
                int su = __builtin_amdgcn_readfirstlane(s0);
                r = shade_entry(scene_at(B, su), render_at(B, su), P, i, slot);
            }
            int pos = wave_append(r.alive, (uint32_t *)&TPT_SEG(B, nb + 1, s0));
            if (r.alive) next[TPT_SEG(B, nb, s0) + pos] = slot;
            int spos = wave_append(r.emit_shadow, (uint32_t *)&TPT_SEG(B, kSegShCnt, s0));
            if (r.emit_shadow) {
                int k = TPT_SEG(B, kSegShBase, s0) + spos;
                P.shadow_slot[k] = slot;
                P.shadow_org[k] = make_float4(r.s_org.x, r.s_org.y, r.s_org.z, r.s_tmax);
                P.shadow_dir[k] = make_float4(r.s_dir.x, r.s_dir.y, r.s_dir.z, 0.f);
                P.shadow_val[k] = make_float4(r.s_val.x, r.s_val.y, r.s_val.z, 0.f);
            }
            todo &= ~__ballot(mine);
        }
    }
}

template <int NL, bool LDS>
__global__ TPT_TRAVERSAL_ATTR void k_shadow_batch(DBatch B, DPaths P) {
    const int start = batch_start_scene(B);
    uint32_t nodes = 0, tris = 0;
    for (int k = 0; k < B.n; k++) {
        int s = start + k < B.n ? start + k : start + k - B.n;
        s = __builtin_amdgcn_readfirstlane(s);
        int cnt = TPT_SEG(B, kSegShCnt, s);
        if (cnt == 0) continue;
        int base = TPT_SEG(B, kSegShBase, s);
        SrcShadow src{P.shadow_org + base, P.shadow_dir + base, P.shadow_val + base, P.shadow_slot + base, P.color};
        FetchGlobal f{(uint32_t *)&TPT_SEG(B, kSegShFetch, s)};
        persistent_trace<true, false, NL, true, LDS>(scene_at(B, s), src, (uint32_t)cnt, f, nodes, tris, 0);
    }
}

// One-ray-per-lane batch traversal: lane i of the grid takes queue position i;
// the scene of position i is the segment holding it (segment bases in
// `sbase`, counts in `scnt`).  A wavefront runs one pass per distinct scene
// among its 64 entries (one pass except where segments meet), so the scene is
// uniform inside a pass and its record is read through scalar loads.
template <bool ANY, int NL, bool LDS, class Body>
__device__ __forceinline__ void batch_lane_rays(const DBatch &B, const int32_t *sbase, const int32_t *scnt, Body body) {
    for (uint32_t base = blockIdx.x * blockDim.x; base < (uint32_t)B.cap; base += gridDim.x * blockDim.x) {
        uint32_t i = base + threadIdx.x;
        int s = 0;
        bool valid = false;
        if (i < (uint32_t)B.cap) {
            s = find_seg(sbase, B.n, (int32_t)i);
            valid = (int32_t)i < sbase[s] + scnt[s];
        }
        uint64_t todo = __ballot(valid);
        while (todo) {
            int lead = __ffsll((unsigned long long)todo) - 1;
            int s0 = __builtin_amdgcn_readlane(s, lead);
            bool mine = valid && s == s0;
            if (mine) body(scene_at(B, __builtin_amdgcn_readfirstlane(s0)), i);
            todo &= ~__ballot(mine);
        }
    }
}

template <int NL, bool LDS>
__global__ void k_extend_batch_lane(DBatch B, DPaths P, int qsel) {
    const int fb = qsel ? kSegQbBase : kSegQaBase;
    const int32_t *queue = qsel ? P.queue_b : P.queue_a;
    batch_lane_rays<false, NL, LDS>(B, B.seg + fb * B.seg_stride, B.seg + (fb + 1) * B.seg_stride,
                                    [&](const DScene &S, uint32_t i) {
                                        uint32_t nodes = 0, tris = 0;
                                        int slot = queue[i];
                                        float4 o4 = P.org[slot], d4 = P.dir[slot];
                                        AlphaCtx ac{P.key[slot], P.depth[slot], true};
                                        HitRec h = trace<false, false, NL, LDS>(S, mk4(o4), mk4(d4), kINF, ac, nodes,
                                                                                tris, 0);
                                        P.hit[i] = make_float4(h.t, __int_as_float(h.loc), h.u, h.v);
                                    });
}

template <int NL, bool LDS>
__global__ void k_shadow_batch_lane(DBatch B, DPaths P) {
    batch_lane_rays<true, NL, LDS>(B, B.seg + kSegShBase * B.seg_stride, B.seg + kSegShCnt * B.seg_stride,
                                   [&](const DScene &S, uint32_t i) {
                                       uint32_t nodes = 0, tris = 0;
                                       AlphaCtx ac{0, 0, false};
                                       float4 o4 = P.shadow_org[i], d4 = P.shadow_dir[i];
                                       HitRec h = trace<true, false, NL, LDS>(S, mk4(o4), mk4(d4), o4.w, ac, nodes,
                                                                              tris, 0);
                                       if (h.gid < 0) {
                                           int slot = P.shadow_slot[i];
                                           float4 c = P.color[slot], v = P.shadow_val[i];
                                           P.color[slot] = make_float4(c.x + v.x, c.y + v.y, c.z + v.z, 0.f);
                                       }
                                   });
}

// Chunks of the scenes' segments for the chunked traversal kernels: the chunk
// size ck = chunk_of(total rays, chunk, waves) goes to out[n + 1]; out[s] =
// exclusive prefix over scenes of ceil(count_s / ck), out[n] = the total.  One
// wavefront.
__global__ void k_batch_chunks(DBatch B, int fcnt, int fout, int chunk, int waves) {
    const int total = wave_scan(B.n, [&](int s) { return TPT_SEG(B, fcnt, s); }, [](int, int) {});
    const int ck = (int)chunk_of((uint32_t)total, chunk, waves);
    const int nch = wave_scan(B.n, [&](int s) { return (TPT_SEG(B, fcnt, s) + ck - 1) / ck; },
                              [&](int s, int excl) { TPT_SEG(B, fout, s) = excl; });
    if (__lane_id() == 0) {
        TPT_SEG(B, fout, B.n) = nch;
        TPT_SEG(B, fout, B.n + 1) = ck;
    }
}

// Chunked batch traversal: wave w takes chunk w of the concatenated per-scene
// chunk lists, i.e. `chunk` consecutive entries of one scene's segment, and
// refills its idle lanes from them; the scene is uniform across the wave.
template <bool ANY, int NL, bool POST, bool LDS, class MakeSrc>
__device__ __forceinline__ void batch_chunk_rays(const DBatch &B, int fbase, int fcnt, int fchunk, MakeSrc make_src) {
    const int32_t *pre = B.seg + fchunk * B.seg_stride;
    int w = __builtin_amdgcn_readfirstlane((int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6));
    if (w >= pre[B.n]) return;
    const uint32_t ck = (uint32_t)pre[B.n + 1];
    int s = __builtin_amdgcn_readfirstlane(find_seg(pre, B.n, w));
    uint32_t base = (uint32_t)TPT_SEG(B, fbase, s), cnt = (uint32_t)TPT_SEG(B, fcnt, s);
    uint32_t begin = base + (uint32_t)(w - pre[s]) * ck;
    uint32_t end = min(begin + ck, base + cnt);
    uint32_t nodes = 0, tris = 0;
    FetchLocal f{begin};
    persistent_trace<ANY, false, NL, POST, LDS>(scene_at(B, s), make_src(), end, f, nodes, tris, 0);
}

template <int NL, bool POST, bool LDS>
__global__ TPT_CHUNK_ATTR void k_extend_batch_chunk(DBatch B, DPaths P, int qsel) {
    const int fb = qsel ? kSegQbBase : kSegQaBase;
    const int32_t *queue = qsel ? P.queue_b : P.queue_a;
    batch_chunk_rays<false, NL, POST, LDS>(B, fb, fb + 1, kSegRayChunk, [&]() {
        return SrcQueue{queue, P.org, P.dir, P.key, P.depth, P.hit};
    });
}

template <int NL, bool POST, bool LDS>
__global__ TPT_CHUNK_ATTR void k_shadow_batch_chunk(DBatch B, DPaths P) {
    batch_chunk_rays<true, NL, POST, LDS>(B, kSegShBase, kSegShCnt, kSegShChunk, [&]() {
        return SrcShadow{P.shadow_org, P.shadow_dir, P.shadow_val, P.shadow_slot, P.color};
    });
}

// One thread per pixel of every scene: adds the samples of this pipeline wave
// that belong to the pixel, in sample order.
__global__ void k_accumulate_batch(DBatch B, DPaths P) {
    uint64_t q = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (q >= B.pix_off[B.n]) return;
    int s = find_seg(B.pix_off, B.n, q);
    const DRender &R = B.renders[s];
    uint64_t p = q - B.pix_off[s];
    uint64_t g0 = B.state[0], nw = batch_wave_paths(B);
    uint64_t a = max(B.path_off[s], g0), b = min(B.path_off[s + 1], g0 + nw);
    if (a >= b) return;
    a -= B.path_off[s];
    b -= B.path_off[s];
    uint64_t px = (uint64_t)R.pixels;
    uint64_t s_lo = a > p ? (a - p + px - 1) / px : 0;
    if (s_lo * px + p >= b) return;
    uint64_t s_hi = (b - 1 - p) / px;
    uint64_t slot0 = B.path_off[s] - g0 + p; // slot of sample 0 of this pixel (may wrap below 0)
    accum_run(R, P, q, (int)s_lo, (int)(s_hi - s_lo + 1), [&](int j) { return slot0 + (s_lo + j) * px; }, false);
}

// Start of a pool's first pipeline wave (`g0`) and the batch's path count.
__global__ void k_batch_state(DBatch B, uint64_t g0, uint64_t total) {
    if (threadIdx.x == 0) {
        B.state[0] = g0;
        B.state[1] = total;
    }
}

// Device wall-clock stamp (constant rate, hipDeviceAttributeWallClockRate) of the
// point in its stream where it runs: completion times inside a captured graph.
__global__ void k_stamp(uint64_t *dst) {
    if (threadIdx.x == 0) *dst = (uint64_t)wall_clock64();
}

__global__ void k_batch_end(DBatch B) {
    if (threadIdx.x == 0) B.state[0] += (uint64_t)B.cap * (uint64_t)B.pools;
}

__global__ void k_batch_zero(DBatch B, DPaths P, int fp64) {
    uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (i >= 3 * B.pix_off[B.n]) return;
    P.sub_f[i] = 0.f;
    P.pix_f[i] = 0.f;
    if (fp64) {
        P.sub_d[i] = 0.0;
        P.pix_d[i] = 0.0;
    }
}

__global__ void k_finalize_batch(DBatch B, DPaths P, float *out) {
    uint64_t q = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (q >= B.pix_off[B.n]) return;
    int s = find_seg(B.pix_off, B.n, q);
    for (int c = 0; c < 3; c++) out[3 * q + c] = final_value(B.renders[s], P, q, c);
}

// ---------------------------------------------------------------- host-visible launchers
static inline size_t lds_bytes(int block) { return (size_t)kLdsRow * (size_t)block * sizeof(int32_t); }

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
void launch_megapath(const DScene &S, const DRender &R, const DPaths &P, int wave_base, int n_paths,
                     uint64_t base_mixed, int grid, int block, int refill_lanes, hipStream_t st) {
    if (refill_lanes > 0)
        hipLaunchKernelGGL(k_megapath<true>, dim3(grid), dim3(block), 0, st, S, R, P, wave_base, n_paths, base_mixed,
                           refill_lanes);
    else
        hipLaunchKernelGGL(k_megapath<false>, dim3(grid), dim3(block), 0, st, S, R, P, wave_base, n_paths, base_mixed,
                           0);
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
// LDS plan of a one-ray-per-lane launch of `block` threads: the stack's first
// kLdsStack entries per lane (rows of kLdsRow words) when S.lds_stack is on and
// they fit, then as much of the node cache (quantized layout) as the rest of the
// workgroup's LDS holds.
struct LaneLds {
    bool stack;
    int cache;
    size_t bytes;
};
static LaneLds lane_lds(const DScene &S, int nl, int block) {
    LaneLds r{S.lds_stack != 0, nl == kNodeBvh4Q ? S.lds_nodes : 0, 0};
    size_t stack = r.stack ? lds_bytes(block) : 0;
    if (stack > (size_t)S.max_lds) {
        r.stack = false;
        stack = 0;
    }
    r.cache = std::max(0, std::min(r.cache, (int)(((size_t)S.max_lds - stack) / 64)));
    r.bytes = stack + (size_t)r.cache * 64;
    return r;
}

static size_t trav_shared(const DScene &S, int nl, bool lds_stack) {
    size_t nodes = nl == kNodeBvh4Q ? (size_t)S.lds_nodes * 64 : 0;
    return nodes + (lds_stack ? lds_bytes(S.trav_block) : 0);
}

// Chunked launches spread their rays over at least 16 wavefronts per CU
// (chunk_of on the device); the grid covers max(max_rays / chunk, that many)
// wavefronts, plus `extra` (one partial chunk per scene of a batch).
static int chunk_waves(int num_cu) { return num_cu * 16; }
static int chunk_grid(uint64_t max_rays, int chunk, int block, int num_cu, uint64_t extra = 0) {
    uint64_t waves = std::max<uint64_t>((max_rays + (uint64_t)chunk - 1) / (uint64_t)chunk, (uint64_t)chunk_waves(num_cu));
    waves += extra;
    uint64_t per_block = (uint64_t)block / 64;
    return (int)std::max<uint64_t>(1, (waves + per_block - 1) / per_block);
}

void launch_extend(const DScene &S, const DPaths &P, const int32_t *queue, int qsel, int nsel, int grid, int block,
                   bool count, uint64_t max_rays, hipStream_t st) {
    by_layout(S.layout, [&](auto L) {
        by_bool(count, [&](auto C) {
            constexpr int NL = decltype(L)::value;
            constexpr bool CNT = decltype(C)::value;
            if (!S.persistent) {
                LaneLds m = lane_lds(S, NL, block);
                by_bool(m.stack, [&](auto D) {
                    hipLaunchKernelGGL((k_extend<CNT, NL, decltype(D)::value>), dim3(grid), dim3(block), m.bytes, st, S,
                                       P, queue, qsel, nsel, m.cache);
                });
                return;
            }
            if (S.persistent >= 3) {
                LaneLds m = lane_lds(S, NL, block);
                int g = chunk_grid(max_rays, S.chunk, block, S.num_cu);
                by_bool(S.persistent == 4, [&](auto Q) {
                    by_bool(m.stack, [&](auto D) {
                        hipLaunchKernelGGL((k_extend_chunk<CNT, NL, decltype(Q)::value, decltype(D)::value>), dim3(g),
                                           dim3(block), m.bytes, st, S, P, queue, qsel, nsel, S.chunk,
                                           chunk_waves(S.num_cu), m.cache);
                    });
                });
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
                  int nsel, int grid, int block, hipStream_t st, const DCache *C) {
    if (C)
        hipLaunchKernelGGL(k_shade<true>, dim3(grid), dim3(block), 0, st, S, R, P, queue, next, qsel, nsel, *C);
    else
        hipLaunchKernelGGL(k_shade<false>, dim3(grid), dim3(block), 0, st, S, R, P, queue, next, qsel, nsel, DCache{});
}
void launch_shadow(const DScene &S, const DPaths &P, int grid, int block, bool count, uint64_t max_rays,
                   hipStream_t st) {
    by_layout(S.layout, [&](auto L) {
        by_bool(count, [&](auto C) {
            constexpr int NL = decltype(L)::value;
            constexpr bool CNT = decltype(C)::value;
            if (!S.persistent) {
                LaneLds m = lane_lds(S, NL, block);
                by_bool(m.stack, [&](auto D) {
                    hipLaunchKernelGGL((k_shadow<CNT, NL, decltype(D)::value>), dim3(grid), dim3(block), m.bytes, st,
                                       S, P, m.cache);
                });
                return;
            }
            if (S.persistent >= 3) {
                LaneLds m = lane_lds(S, NL, block);
                int g = chunk_grid(max_rays, S.chunk, block, S.num_cu);
                by_bool(S.persistent == 4, [&](auto Q) {
                    by_bool(m.stack, [&](auto D) {
                        hipLaunchKernelGGL((k_shadow_chunk<CNT, NL, decltype(Q)::value, decltype(D)::value>), dim3(g),
                                           dim3(block), m.bytes, st, S, P, S.chunk, chunk_waves(S.num_cu), m.cache);
                    });
                });
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
                LaneLds m = lane_lds(S, NL, block);
                by_bool(m.stack, [&](auto D) {
                    hipLaunchKernelGGL((k_trace_closest<CNT, NL, decltype(D)::value>), dim3(grid), dim3(block),
                                       m.bytes, st, S, rays, n, prim, t, counters, m.cache);
                });
                return;
            }
            SrcRays src{rays, prim, t, nullptr};
            if (S.persistent >= 3) {
                LaneLds m = lane_lds(S, NL, block);
                int g = chunk_grid(n, S.chunk, block, S.num_cu);
                by_bool(S.persistent == 4, [&](auto Q) {
                    by_bool(m.stack, [&](auto D) {
                        hipLaunchKernelGGL((k_trace_chunk<false, CNT, NL, decltype(Q)::value, decltype(D)::value>),
                                           dim3(g), dim3(block), m.bytes, st, S, src, (uint32_t)n, counters, S.chunk,
                                           chunk_waves(S.num_cu), m.cache);
                    });
                });
                return;
            }
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
            LaneLds m = lane_lds(S, NL, block);
            by_bool(m.stack, [&](auto D) {
                hipLaunchKernelGGL((k_trace_occluded<NL, decltype(D)::value>), dim3(grid), dim3(block), m.bytes, st, S,
                                   rays, n, occ, m.cache);
            });
            return;
        }
        SrcRays src{rays, nullptr, nullptr, occ};
        if (S.persistent >= 3) {
            LaneLds m = lane_lds(S, NL, block);
            int g = chunk_grid(n, S.chunk, block, S.num_cu);
            by_bool(S.persistent == 4, [&](auto Q) {
                by_bool(m.stack, [&](auto D) {
                    hipLaunchKernelGGL((k_trace_chunk<true, false, NL, decltype(Q)::value, decltype(D)::value>), dim3(g),
                                       dim3(block), m.bytes, st, S, src, (uint32_t)n, counters, S.chunk,
                                       chunk_waves(S.num_cu), m.cache);
                });
            });
            return;
        }
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

// One pipeline wave of a batch (global paths [g0, g0 + cap), g0 read on the
// device): every launch has a fixed shape and all wave-dependent values live in
// device memory, so the sequence can be captured once in a hipGraph and replayed
// ceil(total / cap) times.
void launch_batch_wave(const DBatch &B, const DPaths &P, int max_depth, int layout, int block, int trav_block,
                       bool lds_stack, int mode, int chunk, int max_lds, int num_cu, uint64_t total_pixels,
                       hipStream_t st) {
    const bool lane = mode == 0;
    auto cu_grid = [&](const void *k, int tb, size_t sh) {
        int per_cu = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, tb, sh) != hipSuccess || per_cu < 1) per_cu = 1;
        return num_cu * per_cu;
    };
    int gen_grid = (B.cap + block - 1) / block;
    // the captured shade launch has one shape for every bounce, so it is capped
    // (8192 threads per CU) rather than sized to the wave: late bounces would
    // dispatch mostly empty blocks (256 small rooms: 2048 / 8192 / uncapped
    // threads per CU give 512 / 523 / 498 Mpaths/s)
    int shade_grid = std::min(gen_grid, num_cu * 8192 / block);
    hipLaunchKernelGGL(k_batch_begin, dim3(1), dim3(64), 0, st, B);
    hipLaunchKernelGGL(k_generate_batch, dim3(gen_grid), dim3(block), 0, st, B, P);
    if (mode >= 3) {
        // chunked traversal with the pipeline's block size: at most cap / chunk + n chunks
        bool stk = lds_stack && lds_bytes(block) <= (size_t)max_lds;
        size_t sh = stk ? lds_bytes(block) : 0;
        int cg = chunk_grid((uint64_t)B.cap, chunk, block, num_cu, (uint64_t)B.n);
        by_layout(layout, [&](auto L) {
            by_bool(mode == 4, [&](auto Q) {
                by_bool(stk, [&](auto K) {
                    constexpr int NL = decltype(L)::value;
                    constexpr bool POST = decltype(Q)::value, LDS = decltype(K)::value;
                    for (int b = 0; b <= max_depth; b++) {
                        int qsel = b & 1;
                        int fb = qsel ? kSegQbBase : kSegQaBase;
                        hipLaunchKernelGGL(k_batch_chunks, dim3(1), dim3(64), 0, st, B, fb + 1, (int)kSegRayChunk,
                                           chunk, chunk_waves(num_cu));
                        hipLaunchKernelGGL((k_extend_batch_chunk<NL, POST, LDS>), dim3(cg), dim3(block), sh, st, B, P,
                                           qsel);
                        hipLaunchKernelGGL(k_batch_seg, dim3(1), dim3(64), 0, st, B, qsel);
                        hipLaunchKernelGGL(k_shade_batch, dim3(shade_grid), dim3(block), 0, st, B, P, qsel);
                        hipLaunchKernelGGL(k_batch_chunks, dim3(1), dim3(64), 0, st, B, (int)kSegShCnt,
                                           (int)kSegShChunk, chunk, chunk_waves(num_cu));
                        hipLaunchKernelGGL((k_shadow_batch_chunk<NL, POST, LDS>), dim3(cg), dim3(block), sh, st, B,
                                           P);
                    }
                });
            });
        });
        return;
    }
    if (lane) {
        // one ray per lane: the traversal kernels run with the pipeline's block size
        bool stk = lds_stack && lds_bytes(block) <= (size_t)max_lds;
        size_t sh = stk ? lds_bytes(block) : 0;
        by_layout(layout, [&](auto L) {
            by_bool(stk, [&](auto K) {
                constexpr int NL = decltype(L)::value;
                constexpr bool LDS = decltype(K)::value;
                for (int b = 0; b <= max_depth; b++) {
                    int qsel = b & 1;
                    hipLaunchKernelGGL((k_extend_batch_lane<NL, LDS>), dim3(shade_grid), dim3(block), sh, st, B, P,
                                       qsel);
                    hipLaunchKernelGGL(k_batch_seg, dim3(1), dim3(64), 0, st, B, qsel);
                    hipLaunchKernelGGL(k_shade_batch, dim3(shade_grid), dim3(block), 0, st, B, P, qsel);
                    hipLaunchKernelGGL((k_shadow_batch_lane<NL, LDS>), dim3(shade_grid), dim3(block), sh, st, B, P);
                }
            });
        });
        return;
    }
    by_layout(layout, [&](auto L) {
        by_bool(lds_stack, [&](auto K) {
            constexpr int NL = decltype(L)::value;
            constexpr bool LDS = decltype(K)::value;
            // per-lane stack slots only: the wavefronts of a workgroup may trace different scenes
            size_t sh = LDS ? lds_bytes(trav_block) : 0;
            int eg = cu_grid((const void *)(k_extend_batch<NL, LDS>), trav_block, sh);
            int sg = cu_grid((const void *)(k_shadow_batch<NL, LDS>), trav_block, sh);
            for (int b = 0; b <= max_depth; b++) {
                int qsel = b & 1;
                hipLaunchKernelGGL((k_extend_batch<NL, LDS>), dim3(eg), dim3(trav_block), sh, st, B, P, qsel);
                hipLaunchKernelGGL(k_batch_seg, dim3(1), dim3(64), 0, st, B, qsel);
                hipLaunchKernelGGL(k_shade_batch, dim3(shade_grid), dim3(block), 0, st, B, P, qsel);
                hipLaunchKernelGGL((k_shadow_batch<NL, LDS>), dim3(sg), dim3(trav_block), sh, st, B, P);
            }
        });
    });
}
// The end of a pipeline wave: its samples added to the per-pixel sums (in sample
// order), then the pool's next wave start.
void launch_batch_tail(const DBatch &B, const DPaths &P, uint64_t total_pixels, hipStream_t st) {
    hipLaunchKernelGGL(k_accumulate_batch, dim3((unsigned)((total_pixels + 255) / 256)), dim3(256), 0, st, B, P);
    hipLaunchKernelGGL(k_batch_end, dim3(1), dim3(64), 0, st, B);
}
void launch_batch_state(const DBatch &B, uint64_t g0, uint64_t total, hipStream_t st) {
    hipLaunchKernelGGL(k_batch_state, dim3(1), dim3(64), 0, st, B, g0, total);
}
void launch_stamp(uint64_t *dst, hipStream_t st) { hipLaunchKernelGGL(k_stamp, dim3(1), dim3(64), 0, st, dst); }
void launch_batch_zero(const DBatch &B, const DPaths &P, int fp64, uint64_t total_pixels, hipStream_t st) {
    hipLaunchKernelGGL(k_batch_zero, dim3((unsigned)((3 * total_pixels + 255) / 256)), dim3(256), 0, st, B, P, fp64);
}
void launch_batch_finalize(const DBatch &B, const DPaths &P, float *out, uint64_t total_pixels, hipStream_t st) {
    hipLaunchKernelGGL(k_finalize_batch, dim3((unsigned)((total_pixels + 255) / 256)), dim3(256), 0, st, B, P, out);
}

void launch_fill_u32(uint32_t *p, int n, uint32_t v, hipStream_t st) {
    if (n > 0) hipLaunchKernelGGL(k_fill_u32, dim3((n + 255) / 256), dim3(256), 0, st, p, n, v);
}
void launch_cache_reset(const DCache &C, hipStream_t st) { hipLaunchKernelGGL(k_cache_reset, dim3(1), dim3(64), 0, st, C); }
void launch_cache_apply(const DPaths &P, const DCache &C, const float4 *pred, int n_max, hipStream_t st) {
    hipLaunchKernelGGL(k_cache_apply, dim3((n_max + 255) / 256), dim3(256), 0, st, P, C, pred, n_max);
}
void launch_cache_targets(const DPaths &P, const DCache &C, float4 *target, hipStream_t st) {
    hipLaunchKernelGGL(k_cache_targets, dim3((C.train_cap + 255) / 256), dim3(256), 0, st, P, C, target);
}
void launch_frame_accumulate(const DPaths &P, int pixels, int fp64, int first, int frames, hipStream_t st) {
    hipLaunchKernelGGL(k_frame_accumulate, dim3((pixels + 255) / 256), dim3(256), 0, st, P, pixels, fp64, first, frames);
}
void launch_frame_finalize(const DPaths &P, int pixels, int fp64, int frames, float *out, hipStream_t st) {
    hipLaunchKernelGGL(k_frame_finalize, dim3((pixels + 255) / 256), dim3(256), 0, st, P, pixels, fp64, frames, out);
}

} // namespace tpt
