// scene_build.cpp — flattens a tpt_scene_desc into device-arena sections and
// builds the BVH the gfx950 kernels traverse.
//
// Replaces cuda::Scene(const cpu::Scene&) (li-plus/tinypt src/tinypt/cuda/scene.h:95-174),
// which mirrored every object with its own cudaMalloc and rebuilt one heap-indexed
// median-split BVHTree(faces, 4) per mesh (cuda/object.h:338-374, 478-540).  Here all
// triangles of all meshes share ONE binned-SAH BVH, collapsed to 4-wide nodes of 64
// bytes with 8-bit quantized child boxes (or kept binary / f32 4-wide for A/B runs);
// analytic spheres / rectangles / circles stay a linear list, as the reference's top
// level is.
#include "scene_build.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <numeric>

namespace tpt {

namespace {

[[noreturn]] void fail(int code, const std::string &m) { throw tpt_error(code, m); }

inline float dot3(const float *a, const float *b) { return a[0] * b[0] + (a[1] * b[1] + a[2] * b[2]); }

DTex to_dtex(const tpt_texture &t, int num_images, int need_channels, const std::vector<DImage> &imgs) {
    DTex d;
    d.image = t.image;
    d.channel = t.channel;
    d.v[0] = t.value[0];
    d.v[1] = t.value[1];
    d.v[2] = t.value[2];
    if (t.image >= 0) {
        if (t.image >= num_images) fail(TPT_ERR_INVALID_ARGUMENT, "texture image index out of range");
        if (t.channel < 0 || t.channel + need_channels > imgs[t.image].c)
            fail(TPT_ERR_INVALID_ARGUMENT, "texture channel out of range");
    }
    return d;
}

bool is_zero5(const float *v) { return std::fabs(v[0]) <= 1e-5f && std::fabs(v[1]) <= 1e-5f && std::fabs(v[2]) <= 1e-5f; }

// ---------------------------------------------------------------- binned SAH BVH2
struct BTri {
    float lo[3], hi[3], c[3];
};
struct BNode {
    float lo[3], hi[3];
    int left = -1, right = -1;
    int first = 0, count = 0;
};

struct Builder {
    const std::vector<BTri> &tris;
    std::vector<int> idx;
    std::vector<BNode> nodes;
    int max_depth = 0;
    static constexpr int kBins = 16;
    static constexpr int kLeafTarget = 4;    // largest leaf the SAH may keep (<= kMaxLeaf)
    static constexpr float kNodeCost = 1.f;  // a binary node's traversal cost in triangle tests
    static constexpr int kSahDepth = 32; // below: object-median splits (keeps depth < kStackSize)

    explicit Builder(const std::vector<BTri> &t) : tris(t), idx(t.size()) { std::iota(idx.begin(), idx.end(), 0); }

    static float area(const float *lo, const float *hi) {
        float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
        if (dx < 0 || dy < 0 || dz < 0) return 0.f;
        return 2.f * (dx * dy + dy * dz + dz * dx);
    }

    int build(int b, int e, int depth) {
        int ni = (int)nodes.size();
        nodes.emplace_back();
        max_depth = std::max(max_depth, depth);
        float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
        float clo[3] = {INFINITY, INFINITY, INFINITY}, chi[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (int i = b; i < e; i++) {
            const BTri &t = tris[idx[i]];
            for (int a = 0; a < 3; a++) {
                lo[a] = std::min(lo[a], t.lo[a]);
                hi[a] = std::max(hi[a], t.hi[a]);
                clo[a] = std::min(clo[a], t.c[a]);
                chi[a] = std::max(chi[a], t.c[a]);
            }
        }
        std::memcpy(nodes[ni].lo, lo, sizeof(lo));
        std::memcpy(nodes[ni].hi, hi, sizeof(hi));
        int n = e - b;
        auto make_leaf = [&]() {
            nodes[ni].first = b;
            nodes[ni].count = n;
            return ni;
        };
        if (n <= 1) return make_leaf();

        int best_axis = -1, best_bin = -1;
        float best_cost = INFINITY;
        bool use_sah = depth < kSahDepth;
        if (use_sah) {
            for (int a = 0; a < 3; a++) {
                float ext = chi[a] - clo[a];
                if (!(ext > 0.f)) continue;
                int cnt[kBins] = {0};
                float blo[kBins][3], bhi[kBins][3];
                for (int k = 0; k < kBins; k++)
                    for (int q = 0; q < 3; q++) {
                        blo[k][q] = INFINITY;
                        bhi[k][q] = -INFINITY;
                    }
                float scale = (float)kBins / ext;
                for (int i = b; i < e; i++) {
                    const BTri &t = tris[idx[i]];
                    int k = std::min(kBins - 1, (int)((t.c[a] - clo[a]) * scale));
                    cnt[k]++;
                    for (int q = 0; q < 3; q++) {
                        blo[k][q] = std::min(blo[k][q], t.lo[q]);
                        bhi[k][q] = std::max(bhi[k][q], t.hi[q]);
                    }
                }
                float ra[kBins];
                int rc[kBins];
                float rlo[3] = {INFINITY, INFINITY, INFINITY}, rhi[3] = {-INFINITY, -INFINITY, -INFINITY};
                int rn = 0;
                for (int k = kBins - 1; k > 0; k--) {
                    rn += cnt[k];
                    for (int q = 0; q < 3; q++) {
                        rlo[q] = std::min(rlo[q], blo[k][q]);
                        rhi[q] = std::max(rhi[q], bhi[k][q]);
                    }
                    ra[k] = area(rlo, rhi);
                    rc[k] = rn;
                }
                float llo[3] = {INFINITY, INFINITY, INFINITY}, lhi[3] = {-INFINITY, -INFINITY, -INFINITY};
                int ln = 0;
                for (int k = 0; k < kBins - 1; k++) {
                    ln += cnt[k];
                    for (int q = 0; q < 3; q++) {
                        llo[q] = std::min(llo[q], blo[k][q]);
                        lhi[q] = std::max(lhi[q], bhi[k][q]);
                    }
                    if (ln == 0 || rc[k + 1] == 0) continue;
                    float cost = area(llo, lhi) * (float)ln + ra[k + 1] * (float)rc[k + 1];
                    if (cost < best_cost) {
                        best_cost = cost;
                        best_axis = a;
                        best_bin = k;
                    }
                }
            }
        }
        float parent_area = area(lo, hi);
        if (best_axis >= 0) {
            float split_cost = kNodeCost + (parent_area > 0 ? best_cost / parent_area : 0.f);
            if (n <= kLeafTarget && (float)n <= split_cost) return make_leaf();
        } else if (n <= kLeafTarget) {
            return make_leaf();
        }
        int mid;
        if (best_axis >= 0) {
            int a = best_axis;
            float ext = chi[a] - clo[a];
            float scale = (float)kBins / ext;
            auto it = std::partition(idx.begin() + b, idx.begin() + e, [&](int ti) {
                int k = std::min(kBins - 1, (int)((tris[ti].c[a] - clo[a]) * scale));
                return k <= best_bin;
            });
            mid = (int)(it - idx.begin());
            if (mid == b || mid == e) mid = (b + e) / 2;
        } else {
            // object median on the widest centroid axis (or by index when all coincide)
            int a = 0;
            float ext[3] = {chi[0] - clo[0], chi[1] - clo[1], chi[2] - clo[2]};
            if (ext[1] > ext[a]) a = 1;
            if (ext[2] > ext[a]) a = 2;
            mid = (b + e) / 2;
            if (ext[a] > 0.f)
                std::nth_element(idx.begin() + b, idx.begin() + mid, idx.begin() + e,
                                 [&](int x, int y) { return tris[x].c[a] < tris[y].c[a]; });
        }
        int l = build(b, mid, depth + 1);
        int r = build(mid, e, depth + 1);
        nodes[ni].left = l;
        nodes[ni].right = r;
        return ni;
    }
};

// Quantized 4-wide node (kNodeBvh4Q, 64 B = 4 float4):
//   [0] origin.x origin.y origin.z scale.x
//   [1] scale.y  scale.z  qlo_x    qhi_x        (q*: four 8-bit values, byte i = child i)
//   [2] qlo_y    qhi_y    qlo_z    qhi_z
//   [3] child refs 0..3
// A child's box along an axis is [origin + qlo * scale, origin + qhi * scale].
// The scale is a power of two, so q * scale is exact, and every decoded bound lies
// outside the f32 child bound by at least half a quantum.  The kernel computes a
// slab distance as fma(q, scale * inv, (origin - o) * inv); its rounding error is
// about 2^-23 |origin - o| |inv|, so the quantum is kept >= 2^-16 of the scene's
// magnitude `scene_mag` (max |coordinate| of the scene bounds): the half-quantum
// margin then exceeds that error 64 times over for any ray origin within the
// scene's bounds.
void quantize_wide_node(const float lo[3][4], const float hi[3][4], int n, const int32_t ref[4], double scene_mag,
                        float4 out[4]) {
    float origin[3], scale[3];
    uint32_t qlo[3] = {0, 0, 0}, qhi[3] = {0, 0, 0};
    for (int a = 0; a < 3; a++) {
        double plo = lo[a][0], phi = hi[a][0];
        for (int i = 1; i < n; i++) {
            plo = std::min(plo, (double)lo[a][i]);
            phi = std::max(phi, (double)hi[a][i]);
        }
        double want = std::max({(phi - plo) / 250.0, scene_mag * 0x1p-16, 0x1p-100});
        int e;
        std::frexp(want, &e); // want in [2^(e-1), 2^e)
        double s = std::ldexp(1.0, e);
        for (int attempt = 0;; attempt++) {
            if (attempt > 8) fail(TPT_ERR_RUNTIME, "BVH node quantization failed");
            float px = (float)(plo - s);
            if ((double)px > plo - s) px = std::nextafter(px, -INFINITY);
            bool ok = true;
            uint32_t ql = 0, qh = 0;
            for (int i = 0; i < n && ok; i++) {
                double l = std::floor(((double)lo[a][i] - (double)px) / s - 0.5);
                double h = std::ceil(((double)hi[a][i] - (double)px) / s + 0.5);
                if (l < 0.0 || h > 255.0 || l > h) {
                    ok = false;
                    break;
                }
                // decoded bounds are exact in double: px + q * s
                if ((double)px + l * s > (double)lo[a][i] - 0.5 * s || (double)px + h * s < (double)hi[a][i] + 0.5 * s) {
                    ok = false;
                    break;
                }
                ql |= (uint32_t)l << (8 * i);
                qh |= (uint32_t)h << (8 * i);
            }
            if (ok) {
                // empty slots: inverted boxes (lo 255, hi 0), which the ordered slab
                // test misses for any origin within the scene's bounds
                for (int i = n; i < 4; i++) ql |= 255u << (8 * i);
                origin[a] = px;
                scale[a] = (float)s;
                qlo[a] = ql;
                qhi[a] = qh;
                break;
            }
            s *= 2.0;
        }
    }
    auto u = [](uint32_t v) { return __int_as_float_host((int32_t)v); };
    out[0] = make_float4(origin[0], origin[1], origin[2], scale[0]);
    out[1] = make_float4(scale[1], scale[2], u(qlo[0]), u(qhi[0]));
    out[2] = make_float4(u(qlo[1]), u(qhi[1]), u(qlo[2]), u(qhi[2]));
    out[3] = make_float4(__int_as_float_host(ref[0]), __int_as_float_host(ref[1]), __int_as_float_host(ref[2]),
                         __int_as_float_host(ref[3]));
}

} // namespace

void build_host_scene(const tpt_scene_desc &d, HostScene &H) {
    // node layout: quantized 4-wide 64-B nodes by default; TPT_BVH_LAYOUT=bvh4 selects
    // the f32 4-wide 128-B nodes, TPT_BVH_LAYOUT=bvh2 the binary 64-B nodes
    const char *lay = std::getenv("TPT_BVH_LAYOUT");
    H.layout = kNodeBvh4Q;
    if (lay && std::string(lay) == "bvh4") H.layout = kNodeBvh4;
    if (lay && std::string(lay) == "bvh2") H.layout = kNodeBvh2;
    if (d.camera.width <= 0 || d.camera.height <= 0) fail(TPT_ERR_INVALID_ARGUMENT, "camera resolution must be positive");
    if (d.num_objects < 0 || d.num_materials < 0 || d.num_images < 0 || d.num_meshes < 0 || d.num_delta_lights < 0)
        fail(TPT_ERR_INVALID_ARGUMENT, "negative table size");

    // ---- images -> one texel blob (16-B aligned sub-allocations)
    for (int i = 0; i < d.num_images; i++) {
        const tpt_image &im = d.images[i];
        if (im.width <= 0 || im.height <= 0 || im.channels <= 0 || im.channels > 4 || !im.data)
            fail(TPT_ERR_INVALID_ARGUMENT, "invalid image " + std::to_string(i));
        if (im.format != TPT_IMG_F32 && im.format != TPT_IMG_U8) fail(TPT_ERR_INVALID_ARGUMENT, "invalid image format");
        size_t bytes = (size_t)im.width * im.height * im.channels * (im.format == TPT_IMG_U8 ? 1 : 4);
        size_t off = (H.texels.size() + 15) & ~(size_t)15;
        H.texels.resize(off + bytes);
        std::memcpy(H.texels.data() + off, im.data, bytes);
        H.images.push_back(DImage{im.width, im.height, im.channels, im.format, (uint64_t)off});
    }
    if (H.texels.empty()) H.texels.resize(16);

    // ---- materials (cpu/material.h:194-216)
    for (int i = 0; i < d.num_materials; i++) {
        const tpt_material &m = d.materials[i];
        DMaterial g{};
        int nl = (m.combine == TPT_COMBINE_NONE) ? 1 : 2;
        if (m.combine < 0 || m.combine > 2) fail(TPT_ERR_INVALID_ARGUMENT, "invalid BXDF combination");
        if (m.combine == TPT_COMBINE_MIX && !(0 <= m.weight && m.weight <= 1))
            fail(TPT_ERR_INVALID_ARGUMENT, "[MixBXDF] invalid weight " + std::to_string(m.weight));
        for (int k = 0; k < nl; k++) {
            const tpt_bxdf &b = m.lobe[k];
            if (b.type < 0 || b.type > 4) fail(TPT_ERR_UNSUPPORTED, "unsupported BXDF type");
            DLobe &l = g.lobe[k];
            l.type = b.type;
            l.param = b.param;
            l.specular = (b.type == TPT_BXDF_METAL || b.type == TPT_BXDF_DIELECTRIC) ? 1 : 0;
            l.alpha2 = (b.type == TPT_BXDF_GLOSSY_GGX) ? 2 / (b.param + 2) : 0.f; // material.cpp:118
            l.color = to_dtex(b.color, d.num_images, 3, H.images);
        }
        if (nl == 1) g.lobe[1] = g.lobe[0];
        g.combine = m.combine;
        g.weight = m.weight;
        g.specular = g.lobe[0].specular || (nl == 2 && g.lobe[1].specular);
        g.emission = to_dtex(m.emission, d.num_images, 3, H.images);
        g.alpha = to_dtex(m.alpha, d.num_images, 1, H.images);
        g.alpha_opaque = (m.alpha.image < 0 && m.alpha.value[0] == 1.f) ? 1 : 0;
        g.emissive = (m.emission.image >= 0 || !is_zero5(m.emission.value)) ? 1 : 0;
        g.bump = m.bump_image;
        if (g.bump >= 0 && (g.bump >= d.num_images || H.images[g.bump].c != 2 || H.images[g.bump].fmt != TPT_IMG_F32))
            fail(TPT_ERR_INVALID_ARGUMENT, "bump image must be a 2-channel f32 gradient");
        H.materials.push_back(g);
    }
    auto check_mat = [&](int m) {
        if (m < 0 || m >= d.num_materials) fail(TPT_ERR_INVALID_ARGUMENT, "material index out of range");
    };

    // ---- objects: global primitive ids in object order, faces in order
    struct TopRec {
        int type;
        int analytic;            // analytic index
        std::vector<int> shape_first, shape_end; // tri_id ranges per shape
    };
    std::vector<TopRec> top;
    std::vector<BTri> btris;
    int32_t gid = 0;
    int32_t normal_base = 0, tc_base = 0;
    for (int i = 0; i < d.num_objects; i++) {
        const tpt_object &o = d.objects[i];
        TopRec tr{o.type, -1, {}, {}};
        if (o.type == TPT_OBJ_MESH) {
            if (o.mesh < 0 || o.mesh >= d.num_meshes) fail(TPT_ERR_INVALID_ARGUMENT, "mesh index out of range");
            const tpt_mesh &m = d.meshes[o.mesh];
            if (m.num_faces < 0 || m.num_vertices < 0 || m.num_shapes < 0)
                fail(TPT_ERR_INVALID_ARGUMENT, "invalid mesh sizes");
            for (int k = 0; k < m.num_normals * 3; k++) H.normals.push_back(m.normals[k]);
            for (int k = 0; k < m.num_texcoords * 2; k++) H.texcoords.push_back(m.texcoords[k]);
            int tri_base = (int)H.tri_info.size() / 2;
            for (int f = 0; f < m.num_faces; f++) {
                const int32_t *fi = m.faces + (size_t)f * 9;
                for (int c = 0; c < 3; c++) {
                    if (fi[3 * c] < 0 || fi[3 * c] >= m.num_vertices)
                        fail(TPT_ERR_INVALID_ARGUMENT, "face vertex index out of range");
                    if (fi[3 * c + 1] >= m.num_normals || fi[3 * c + 2] >= m.num_texcoords)
                        fail(TPT_ERR_INVALID_ARGUMENT, "face attribute index out of range");
                }
                bool hn = fi[1] >= 0, ht = fi[2] >= 0;
                if ((fi[4] >= 0) != hn || (fi[7] >= 0) != hn || (fi[5] >= 0) != ht || (fi[8] >= 0) != ht)
                    fail(TPT_ERR_INVALID_ARGUMENT, "face corners must agree on normal / texcoord presence");
                int mat = m.face_material[f];
                check_mat(mat);
                const float *va = m.vertices + 3 * fi[0], *vb = m.vertices + 3 * fi[3], *vc = m.vertices + 3 * fi[6];
                float a[3] = {va[0], va[1], va[2]};
                float ab[3] = {vb[0] - va[0], vb[1] - va[1], vb[2] - va[2]}; // object.cpp:205-207
                float ac[3] = {vc[0] - va[0], vc[1] - va[1], vc[2] - va[2]};
                int tri_id = tri_base + f;
                H.tris.push_back(make_float4(a[0], a[1], a[2], __int_as_float_host(gid + f)));
                H.tris.push_back(make_float4(ab[0], ab[1], ab[2], 0.f));
                H.tris.push_back(make_float4(ac[0], ac[1], ac[2], 0.f));
                int na = hn ? fi[1] + normal_base : -1, nb = hn ? fi[4] + normal_base : -1,
                    nc = hn ? fi[7] + normal_base : -1;
                int ta = ht ? fi[2] + tc_base : -1, tb = ht ? fi[5] + tc_base : -1, tc = ht ? fi[8] + tc_base : -1;
                H.tri_info.push_back(make_int4(mat, na, nb, nc));
                H.tri_info.push_back(make_int4(ta, tb, tc, o.mesh));
                BTri bt;
                float b2[3] = {a[0] + ab[0], a[1] + ab[1], a[2] + ab[2]};
                float c2[3] = {a[0] + ac[0], a[1] + ac[1], a[2] + ac[2]};
                for (int q = 0; q < 3; q++) {
                    bt.lo[q] = std::min(std::min(a[q], b2[q]), c2[q]);
                    bt.hi[q] = std::max(std::max(a[q], b2[q]), c2[q]);
                    bt.c[q] = 0.5f * (bt.lo[q] + bt.hi[q]);
                }
                btris.push_back(bt);
                H.gid_local.push_back(tri_id);
                (void)tri_id;
            }
            for (int s = 0; s < m.num_shapes; s++) {
                int f0 = m.shape_offsets[s], f1 = m.shape_offsets[s + 1];
                if (f0 < 0 || f1 < f0 || f1 > m.num_faces) fail(TPT_ERR_INVALID_ARGUMENT, "invalid shape offsets");
                tr.shape_first.push_back(tri_base + f0);
                tr.shape_end.push_back(tri_base + f1);
            }
            gid += m.num_faces;
            normal_base += m.num_normals;
            tc_base += m.num_texcoords;
        } else {
            check_mat(o.material);
            DAnalytic a{};
            a.material = o.material;
            a.gid = gid;
            std::memcpy(a.center, o.center, sizeof(a.center));
            a.radius = o.radius;
            std::memcpy(a.u, o.rot, 3 * sizeof(float));
            std::memcpy(a.v, o.rot + 3, 3 * sizeof(float));
            std::memcpy(a.w, o.rot + 6, 3 * sizeof(float));
            a.dim[0] = o.dim[0];
            a.dim[1] = o.dim[1];
            if (o.type == TPT_OBJ_SPHERE) {
                a.type = A_SPHERE;
            } else if (o.type == TPT_OBJ_RECTANGLE) {
                a.type = A_RECT;
                a.dist = dot3(a.w, a.center); // Rectangle::calc_dist (object.h:116)
            } else if (o.type == TPT_OBJ_CIRCLE) {
                a.type = A_CIRCLE;
                a.dist = dot3(a.w, a.center); // object.cpp:157
            } else {
                fail(TPT_ERR_UNSUPPORTED, "unsupported object type " + std::to_string(o.type));
            }
            tr.analytic = (int)H.analytic.size();
            H.gid_local.push_back(-(tr.analytic + 1));
            H.analytic.push_back(a);
            gid += 1;
        }
        top.push_back(std::move(tr));
    }
    if (H.normals.empty()) H.normals.resize(3, 0.f);
    if (H.texcoords.empty()) H.texcoords.resize(2, 0.f);
    if (H.gid_local.empty()) H.gid_local.push_back(0);

    // ---- BVH over all triangles
    int ntri = (int)btris.size();
    if (ntri > (1 << 27)) fail(TPT_ERR_UNSUPPORTED, "too many triangles for the leaf encoding");
    if (ntri > 0) {
        Builder B(btris);
        B.build(0, ntri, 0);
        H.max_depth_seen = B.max_depth;
        // flatten: inner nodes get consecutive indices (DFS pre-order), leaves become child refs
        std::vector<int> inner_index(B.nodes.size(), -1);
        std::vector<int> order;
        std::vector<int> stk{0};
        while (!stk.empty()) {
            int n = stk.back();
            stk.pop_back();
            if (B.nodes[n].left < 0) continue;
            inner_index[n] = (int)order.size();
            order.push_back(n);
            stk.push_back(B.nodes[n].right);
            stk.push_back(B.nodes[n].left);
        }
        // triangles in leaf order, then one zero-size sentinel triangle: the empty
        // slots of 4-wide nodes are leaves holding it (its test divides by a zero
        // determinant and the NaN barycentrics reject it), so the traversal needs
        // no empty-slot check
        H.tris_bvh.assign((size_t)(ntri + 1) * 3, make_float4(0.f, 0.f, 0.f, 0.f));
        H.tris_bvh[3 * (size_t)ntri] = make_float4(0.f, 0.f, 0.f, __int_as_float_host(INT32_MAX));
        const int32_t sentinel_leaf = (int32_t)(kLeafBit | ((uint32_t)ntri << 3));
        // tri ids need loc_bits bits; the material index goes above them when it fits
        // in the remaining 31 (the sign marks analytic locations)
        {
            int bits = 1;
            while (bits < 31 && ((int64_t)1 << bits) < (int64_t)ntri) bits++;
            H.loc_bits = bits < 31 && (int64_t)H.materials.size() <= ((int64_t)1 << (31 - bits)) ? bits : 0;
        }
        int cursor = 0;
        std::vector<int> leaf_first(B.nodes.size(), -1);
        std::vector<int> stk2{0};
        while (!stk2.empty()) {
            int n = stk2.back();
            stk2.pop_back();
            const BNode &bn = B.nodes[n];
            if (bn.left < 0) {
                leaf_first[n] = cursor;
                for (int k = 0; k < bn.count; k++) {
                    int tri = B.idx[bn.first + k];
                    float4 r0 = H.tris[3 * tri], r1 = H.tris[3 * tri + 1], r2 = H.tris[3 * tri + 2];
                    int mat = H.tri_info[2 * tri].x;
                    int flags = H.materials[mat].alpha_opaque ? 0 : 1;
                    H.tris_bvh[3 * cursor] = r0;
                    const int32_t loc = H.loc_bits ? (int32_t)((uint32_t)tri | (uint32_t)mat << H.loc_bits) : tri;
                    H.tris_bvh[3 * cursor + 1] = make_float4(r1.x, r1.y, r1.z, __int_as_float_host(loc));
                    H.tris_bvh[3 * cursor + 2] = make_float4(r2.x, r2.y, r2.z, __int_as_float_host(flags));
                    cursor++;
                }
                continue;
            }
            stk2.push_back(bn.right);
            stk2.push_back(bn.left);
        }
        auto leaf_ref = [&](int n) -> int32_t {
            const BNode &bn = B.nodes[n];
            if (bn.count < 1 || bn.count > kMaxLeaf) fail(TPT_ERR_RUNTIME, "leaf size out of range");
            return (int32_t)(kLeafBit | ((uint32_t)leaf_first[n] << 3) | (uint32_t)(bn.count - 1));
        };
        if (H.layout == kNodeBvh2) {
            auto ref = [&](int n) -> int32_t { return B.nodes[n].left >= 0 ? inner_index[n] : leaf_ref(n); };
            H.nodes.resize(order.size() * 4);
            for (size_t k = 0; k < order.size(); k++) {
                const BNode &bn = B.nodes[order[k]];
                const BNode &L = B.nodes[bn.left], &R = B.nodes[bn.right];
                H.nodes[4 * k + 0] = make_float4(L.lo[0], L.lo[1], L.lo[2], L.hi[0]);
                H.nodes[4 * k + 1] = make_float4(L.hi[1], L.hi[2], R.lo[0], R.lo[1]);
                H.nodes[4 * k + 2] = make_float4(R.lo[2], R.hi[0], R.hi[1], R.hi[2]);
                H.nodes[4 * k + 3] = make_float4(__int_as_float_host(ref(bn.left)), __int_as_float_host(ref(bn.right)), 0.f, 0.f);
            }
            H.root = ref(0);
            H.max_stack = B.max_depth;
        } else {
            // collapse the binary tree into 4-wide nodes: open the inner child of largest
            // surface area until four children are gathered (or only leaves remain)
            struct W {
                int bnode;      // binary node this wide node replaces
                int children[4];
                int n;
            };
            double scene_mag = 0.0;
            for (int a = 0; a < 3; a++)
                scene_mag = std::max({scene_mag, std::fabs((double)B.nodes[0].lo[a]), std::fabs((double)B.nodes[0].hi[a])});
            std::vector<W> wide;
            std::vector<int> wide_of(B.nodes.size(), -1);
            std::vector<int> todo{0};
            while (!todo.empty()) {
                int bn = todo.back();
                todo.pop_back();
                W w{bn, {-1, -1, -1, -1}, 0};
                std::vector<int> ch{B.nodes[bn].left, B.nodes[bn].right};
                while ((int)ch.size() < 4) {
                    int best = -1;
                    float best_area = -1.f;
                    for (int i = 0; i < (int)ch.size(); i++) {
                        const BNode &c = B.nodes[ch[i]];
                        if (c.left < 0) continue;
                        float a = Builder::area(c.lo, c.hi);
                        if (a > best_area) {
                            best_area = a;
                            best = i;
                        }
                    }
                    if (best < 0) break;
                    int open = ch[best];
                    ch.erase(ch.begin() + best);
                    ch.insert(ch.begin() + best, B.nodes[open].right);
                    ch.insert(ch.begin() + best, B.nodes[open].left);
                }
                w.n = (int)ch.size();
                for (int i = 0; i < w.n; i++) w.children[i] = ch[i];
                wide_of[bn] = (int)wide.size();
                wide.push_back(w);
                for (int i = w.n - 1; i >= 0; i--)
                    if (B.nodes[ch[i]].left >= 0) todo.push_back(ch[i]);
            }
            // node order: the first kBfsPrefix nodes in breadth-first order (the top of
            // the tree, which the traversal kernels may stage in LDS), the rest in the
            // depth-first pre-order they were created in
            std::vector<int> pos(wide.size(), -1);
            int next = 0;
            std::vector<int> bfs{0};
            for (size_t qh = 0; qh < bfs.size() && next < kBfsPrefix; qh++) {
                const W &w = wide[bfs[qh]];
                pos[bfs[qh]] = next++;
                for (int i = 0; i < w.n; i++)
                    if (B.nodes[w.children[i]].left >= 0) bfs.push_back(wide_of[w.children[i]]);
            }
            H.bfs_prefix = next;
            for (size_t k = 0; k < wide.size(); k++)
                if (pos[k] < 0) pos[k] = next++;
            auto ref = [&](int n) -> int32_t { return B.nodes[n].left >= 0 ? pos[wide_of[n]] : leaf_ref(n); };
            const int nf4 = kNodeFloat4[H.layout];
            H.nodes.resize(wide.size() * nf4);
            for (size_t wk = 0; wk < wide.size(); wk++) {
                const W &w = wide[wk];
                const size_t k = (size_t)pos[wk];
                float lo[3][4], hi[3][4];
                int32_t r[4];
                for (int i = 0; i < 4; i++) {
                    if (i < w.n) {
                        const BNode &c = B.nodes[w.children[i]];
                        for (int a = 0; a < 3; a++) {
                            lo[a][i] = c.lo[a];
                            hi[a][i] = c.hi[a];
                        }
                        r[i] = ref(w.children[i]);
                    } else {
                        for (int a = 0; a < 3; a++) lo[a][i] = hi[a][i] = 0.f;
                        r[i] = sentinel_leaf;
                    }
                }
                if (H.layout == kNodeBvh4Q) {
                    quantize_wide_node(lo, hi, w.n, r, scene_mag, &H.nodes[4 * k]);
                    continue;
                }
                for (int a = 0; a < 3; a++) {
                    H.nodes[8 * k + 2 * a] = make_float4(lo[a][0], lo[a][1], lo[a][2], lo[a][3]);
                    H.nodes[8 * k + 2 * a + 1] = make_float4(hi[a][0], hi[a][1], hi[a][2], hi[a][3]);
                }
                H.nodes[8 * k + 6] = make_float4(__int_as_float_host(r[0]), __int_as_float_host(r[1]),
                                                 __int_as_float_host(r[2]), __int_as_float_host(r[3]));
                H.nodes[8 * k + 7] = make_float4(0.f, 0.f, 0.f, 0.f);
            }
            H.root = ref(0);
            // worst-case stack: a node pushes (hit children - 1); bound = max over root-to-node
            // paths of the sum of (children - 1)
            int worst = 0;
            std::vector<std::pair<int, int>> st{{0, 0}};
            while (!st.empty()) {
                auto [wi, depth] = st.back();
                st.pop_back();
                const W &w = wide[wi];
                int here = depth + (w.n - 1);
                worst = std::max(worst, here);
                for (int i = 0; i < w.n; i++)
                    if (B.nodes[w.children[i]].left >= 0) st.push_back({wide_of[w.children[i]], here});
            }
            H.max_stack = worst;
        }
        if (H.max_stack >= kStackSize) fail(TPT_ERR_RUNTIME, "BVH worst-case stack exceeds the traversal stack");
    }
    if (H.tris.empty()) {
        H.tris.resize(3, make_float4(0, 0, 0, 0));
        H.tri_info.resize(2, make_int4(0, -1, -1, -1));
    }
    if (H.tris_bvh.empty()) H.tris_bvh.resize(3, make_float4(0, 0, 0, 0));
    // ---- shading records: one 128-B line per triangle with what shading reads
    {
        const size_t nt = H.tri_info.size() / 2;
        H.shade_rec.assign(nt * kShadeRec, make_float4(0, 0, 0, 0));
        for (size_t t = 0; t < nt; t++) {
            const int4 i0 = H.tri_info[2 * t], i1 = H.tri_info[2 * t + 1];
            const float4 ab = H.tris[3 * t + 1], ac = H.tris[3 * t + 2];
            const int flags = (i0.y >= 0 ? 1 : 0) | (i1.x >= 0 ? 2 : 0);
            float n[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, uv[6] = {0, 0, 0, 0, 0, 0};
            if (flags & 1)
                for (int c = 0; c < 3; c++) {
                    int ni = c == 0 ? i0.y : (c == 1 ? i0.z : i0.w);
                    for (int q = 0; q < 3; q++) n[3 * c + q] = H.normals[3 * (size_t)ni + q];
                }
            if (flags & 2)
                for (int c = 0; c < 3; c++) {
                    int ti = c == 0 ? i1.x : (c == 1 ? i1.y : i1.z);
                    for (int q = 0; q < 2; q++) uv[2 * c + q] = H.texcoords[2 * (size_t)ti + q];
                }
            float4 *r = &H.shade_rec[t * kShadeRec];
            r[0] = make_float4(ab.x, ab.y, ab.z, __int_as_float_host(i0.x));
            r[1] = make_float4(ac.x, ac.y, ac.z, __int_as_float_host(flags));
            r[2] = make_float4(n[0], n[1], n[2], n[3]);
            r[3] = make_float4(n[4], n[5], n[6], n[7]);
            r[4] = make_float4(n[8], uv[0], uv[1], uv[2]);
            r[5] = make_float4(uv[3], uv[4], uv[5], 0.f);
        }
    }
    if (H.nodes.empty()) H.nodes.resize(8, make_float4(0, 0, 0, 0));

    // ---- object lights: reference DFS order (cpu/scene.cpp:9-19)
    {
        struct N {
            int kind, a, b;
        };
        std::vector<N> stk{{0, 0, 0}};
        while (!stk.empty()) {
            N n = stk.back();
            stk.pop_back();
            if (n.kind == 0) {
                for (int i = 0; i < (int)top.size(); i++) stk.push_back({1, i, 0});
            } else if (n.kind == 1) {
                const TopRec &t = top[n.a];
                if (t.type == TPT_OBJ_MESH) {
                    for (int s = 0; s < (int)t.shape_first.size(); s++) stk.push_back({2, n.a, s});
                } else {
                    const DAnalytic &a = H.analytic[t.analytic];
                    if (H.materials[a.material].emissive) H.lights.push_back(DLight{L_ANALYTIC, t.analytic, a.gid, 0});
                }
            } else if (n.kind == 2) {
                const TopRec &t = top[n.a];
                for (int tri = t.shape_first[n.b]; tri < t.shape_end[n.b]; tri++) stk.push_back({3, tri, 0});
            } else {
                int tri = n.a;
                if (H.materials[H.tri_info[2 * tri].x].emissive)
                    H.lights.push_back(DLight{L_TRIANGLE, tri, __float_as_int_host(H.tris[3 * tri].w), 0});
            }
        }
    }

    // ---- non-delta lights passed to Scene(), after the emissive objects (cpu/scene.cpp:21-27)
    H.light_shapes.clear();
    for (int i = 0; i < d.num_object_lights; i++) {
        if (!d.object_lights) fail(TPT_ERR_INVALID_ARGUMENT, "object_lights is null");
        const tpt_object_light &ol = d.object_lights[i];
        if (ol.kind == TPT_OBJECT_LIGHT_ENV) {
            H.lights.push_back(DLight{L_ENV, -1, -1, 0});
            continue;
        }
        if (ol.kind != TPT_OBJECT_LIGHT_OBJECT) fail(TPT_ERR_UNSUPPORTED, "unsupported object light kind");
        if (ol.object < 0 || ol.object >= (int)top.size()) fail(TPT_ERR_INVALID_ARGUMENT, "object light index out of range");
        const TopRec &t = top[ol.object];
        if (t.type != TPT_OBJ_MESH) {
            H.lights.push_back(DLight{L_ANALYTIC, t.analytic, H.analytic[t.analytic].gid, 0});
            continue;
        }
        const int first = (int)H.light_shapes.size();
        for (size_t s = 0; s < t.shape_first.size(); s++)
            if (t.shape_end[s] > t.shape_first[s]) H.light_shapes.push_back(int2{t.shape_first[s], t.shape_end[s]});
        const int ns = (int)H.light_shapes.size() - first;
        if (ns == 0) fail(TPT_ERR_INVALID_ARGUMENT, "ObjectLight of a mesh without faces");
        H.lights.push_back(DLight{L_MESH, first, ns, 0});
    }

    // ---- delta lights (cpu/light.h:45-111)
    for (int i = 0; i < d.num_delta_lights; i++) {
        const tpt_delta_light &l = d.delta_lights[i];
        DDelta g{};
        if (l.type != TPT_LIGHT_DISTANT && l.type != TPT_LIGHT_POINT) fail(TPT_ERR_UNSUPPORTED, "unsupported light type");
        g.type = l.type == TPT_LIGHT_DISTANT ? 0 : 1;
        std::memcpy(g.u, l.rot, 3 * sizeof(float));
        std::memcpy(g.v, l.rot + 3, 3 * sizeof(float));
        std::memcpy(g.w, l.rot + 6, 3 * sizeof(float));
        std::memcpy(g.center, l.center, sizeof(g.center));
        for (int c = 0; c < 3; c++) g.emission[c] = l.color[c] * l.power; // emission() = color * power
        g.radius = l.radius;
        if (l.type == TPT_LIGHT_DISTANT) {
            if (!(0 <= l.radius && l.radius <= 3.14159265358979323846f))
                fail(TPT_ERR_INVALID_ARGUMENT, "angle out of range: " + std::to_string(l.radius));
            g.angular_radius = l.radius / 2; // DistantLight ctor (light.cpp:6-9)
        }
        H.delta.push_back(g);
    }

    // ---- background / EnvLight (scene.cpp:25-30; texture.cpp:65-77)
    H.bg = to_dtex(d.background.rgb, d.num_images, 3, H.images);
    std::memcpy(H.bg_rot, d.background.rot, sizeof(H.bg_rot));
    H.bg_map = d.background.rgb.image >= 0 ? 1 : 0;
    auto build_cdf = [](const std::vector<float> &w, std::vector<float> &cdf) { // distribution.cpp:6-30
        size_t n = w.size();
        size_t off = cdf.size();
        cdf.resize(off + n + 1);
        float *c = cdf.data() + off;
        c[0] = 0;
        for (size_t i = 0; i < n; i++) c[i + 1] = c[i] + w[i];
        float sum = c[n];
        if (sum > 0) {
            for (size_t i = 0; i <= n; i++) c[i] /= sum;
        } else {
            for (size_t i = 0; i <= n; i++) c[i] = (float)i / (float)n;
        }
    };
    if (H.bg_map && d.background.importance) {
        const DImage &im = H.images[H.bg.image];
        H.env_w = im.w;
        H.env_h = im.h;
        std::vector<float> rowsum(im.h);
        for (int y = 0; y < im.h; y++) {
            std::vector<float> w(im.w);
            for (int x = 0; x < im.w; x++) {
                float px[3];
                for (int c = 0; c < 3; c++) {
                    size_t idx = ((size_t)y * im.w + x) * im.c + H.bg.channel + c;
                    if (im.fmt == TPT_IMG_U8) px[c] = (float)H.texels[im.offset + idx] / 255.f;
                    else std::memcpy(&px[c], H.texels.data() + im.offset + idx * 4, 4);
                }
                w[x] = std::max(px[0], std::max(px[1], px[2]));
                if (w[x] < 0) fail(TPT_ERR_INVALID_ARGUMENT, "negative weight");
            }
            rowsum[y] = std::accumulate(w.begin(), w.end(), 0.f);
            build_cdf(w, H.env_cond);
        }
        build_cdf(rowsum, H.env_marg);
    } else {
        H.env_w = 1;
        H.env_h = 1;
        H.env_cond = {0.f, 1.f};
        H.env_marg = {0.f, 1.f};
    }
    // guide tables: sampling a CDF of `count` bins with u in bucket k = floor(u * G)
    // needs upper_bound only over [g[k], g[k + 1]] (one or two entries), the same
    // index as the search over the whole CDF
    auto pow2_at_least = [](int n) {
        int g = 1;
        while (g < n) g <<= 1;
        return g;
    };
    auto build_guide = [](const float *cdf, int count, int G, std::vector<int32_t> &out) {
        for (int k = 0; k < G; k++) {
            const float t = (float)k / (float)G; // exact: G is a power of two
            out.push_back((int32_t)(std::upper_bound(cdf, cdf + count + 1, t) - cdf));
        }
        out.push_back(count + 1);
    };
    H.env_gh = pow2_at_least(H.env_h);
    H.env_gw = pow2_at_least(H.env_w);
    H.env_gmarg.clear();
    H.env_gcond.clear();
    build_guide(H.env_marg.data(), H.env_h, H.env_gh, H.env_gmarg);
    for (int y = 0; y < H.env_h; y++)
        build_guide(H.env_cond.data() + (size_t)y * (H.env_w + 1), H.env_w, H.env_gw, H.env_gcond);
    if (H.bg_map) H.lights.push_back(DLight{L_ENV, -1, -1, 0});

    // ---- camera (cpu/camera.cpp:6-21)
    const tpt_camera &c = d.camera;
    DCamera &cam = H.cam;
    std::memcpy(cam.center, c.center, sizeof(cam.center));
    std::memcpy(cam.u, c.rot, 3 * sizeof(float));
    std::memcpy(cam.v, c.rot + 3, 3 * sizeof(float));
    std::memcpy(cam.w, c.rot + 6, 3 * sizeof(float));
    cam.width = c.width;
    cam.height = c.height;
    cam.aperture = c.aperture;
    float vw, vh;
    if (c.width > c.height) {
        vw = 2.f * std::tan(c.fov * .5f);
        vh = vw * (float)c.height / (float)c.width;
    } else {
        vh = 2.f * std::tan(c.fov * .5f);
        vw = vh * (float)c.width / (float)c.height;
    }
    float fr = c.focus_dist * vw, fu = c.focus_dist * vh;
    for (int k = 0; k < 3; k++) {
        cam.direct[k] = c.focus_dist * -cam.w[k];
        cam.right[k] = fr * cam.u[k];
        cam.up[k] = fu * cam.v[k];
    }
}

} // namespace tpt
