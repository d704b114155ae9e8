// oracle.cpp — TEST INFRASTRUCTURE ONLY: the CPU parity oracle.
//
// A restatement of the reference CPU renderer (li-plus/tinypt, src/tinypt/cpu/)
// over the flat tpt_scene_desc.  Each function names the reference code it
// follows.  Deviations that the HIP path shares, fixed here on purpose:
//   * one erand48 stream per (pixel, sample) instead of per row (oracle_rng.h);
//   * the alpha cut-out draw is a hash of (path, depth, primitive) and is only
//     taken for materials whose alpha is not the constant 1, instead of one
//     stream draw per improving candidate (cpu/object.cpp:53,285);
//   * mesh Mapping is baked into world-space vertices by the caller;
//   * unspecified C++ operand orders are fixed left to right.
#include "oracle.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <memory>
#include <numeric>
#include <stdexcept>
#include <string>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

#include "oracle_math.h"
#include "oracle_rng.h"

using namespace oracle;

namespace {

thread_local std::string g_last_error;

// ---------------------------------------------------------------- textures
struct Img {
    int w = 0, h = 0, c = 0, fmt = 0;
    std::vector<uint8_t> bytes;
    float texel(int x, int y, int ch) const {
        size_t idx = ((size_t)y * (size_t)w + (size_t)x) * (size_t)c + (size_t)ch;
        if (fmt == TPT_IMG_U8) return (float)bytes[idx] / 255.f;
        float v;
        std::memcpy(&v, bytes.data() + idx * 4, 4);
        return v;
    }
};

struct Tex {
    int image = -1;
    int channel = 0;
    V3 value{0, 0, 0};
};

struct Scene;

// denormalize_coords (cpu/texture.cpp:6-14)
static inline void denormalize(V2 uv, int w, int h, int &x, int &y) {
    float iu = uv.x - std::floor(uv.x);
    float iv = uv.y - std::floor(uv.y);
    x = (int)(iu * ((float)w - EPS));
    y = (int)(iv * ((float)h - EPS));
}

// ---------------------------------------------------------------- distributions (cpu/distribution.cpp)
struct Dist1D {
    std::vector<float> cdf;
    Dist1D() : Dist1D(std::vector<float>{1.f}) {}
    explicit Dist1D(const std::vector<float> &w) { // distribution.cpp:6-30
        for (float x : w)
            if (x < 0) throw std::invalid_argument("negative weight");
        if (w.empty()) {
            cdf = {0.f, 1.f};
            return;
        }
        cdf.resize(w.size() + 1);
        cdf[0] = 0;
        for (size_t i = 0; i < w.size(); i++) cdf[i + 1] = cdf[i] + w[i];
        float sum = cdf.back();
        if (sum > 0) {
            for (auto &v : cdf) v /= sum;
        } else {
            for (size_t i = 0; i < cdf.size(); i++) cdf[i] = (float)i / (float)count();
        }
    }
    int count() const { return (int)cdf.size() - 1; }
    float sample(float u) const { // distribution.cpp:32-39
        int i = (int)(std::upper_bound(cdf.begin(), cdf.end(), u) - cdf.begin());
        i = std::min(i, count());
        return ((float)i - 1 + (u - cdf[i - 1]) / (cdf[i] - cdf[i - 1])) / (float)count();
    }
    int to_anchor(float x) const { return (int)(x * ((float)count() - EPS)); } // distribution.h:23
    float pdf(float x) const {                                                 // distribution.cpp:41-44
        int i = to_anchor(x);
        return (cdf[i + 1] - cdf[i]) * (float)count();
    }
};

struct Dist2D { // distribution.cpp:55-90
    std::vector<Dist1D> cond;
    Dist1D marg;
    Dist2D() : Dist2D(std::vector<std::vector<float>>{{1.f}}) {}
    explicit Dist2D(const std::vector<std::vector<float>> &w) {
        if (w.empty() || w.front().empty()) throw std::invalid_argument("empty weights");
        size_t xs = w.front().size();
        std::vector<float> yw;
        for (const auto &row : w) {
            if (row.size() != xs) throw std::invalid_argument("irregular weights");
            yw.push_back(std::accumulate(row.begin(), row.end(), 0.f));
            cond.emplace_back(row);
        }
        marg = Dist1D(yw);
    }
    V2 sample(float u, float v) const {
        float y = marg.sample(v);
        int yi = marg.to_anchor(y);
        float x = cond[yi].sample(u);
        return {x, y};
    }
    float pdf(V2 xy) const {
        float py = marg.pdf(xy.y);
        int yi = marg.to_anchor(xy.y);
        return py * cond[yi].pdf(xy.x);
    }
};

// EnvTexture::uv2dir / dir2uv (cpu/texture.cpp:102-121)
static V3 uv2dir(V2 uv) {
    float u = uv.x, v = uv.y;
    float phi = 2 * PI_F * (1 - u) - PI_F;
    float theta = PI_F * (1 - v);
    float cos_theta = std::cos(theta);
    float sin_theta = std::sqrt(1 - cos_theta * cos_theta);
    return {std::cos(phi) * sin_theta, std::sin(phi) * sin_theta, cos_theta};
}
static V2 dir2uv(V3 d) {
    float u = 1 - (std::atan2(d.y, d.x) + PI_F) * (.5f * INV_PI_F);
    float v = 1 - std::acos(std::clamp(d.z, -1.f, 1.f)) * INV_PI_F;
    return {u, v};
}

// ---------------------------------------------------------------- materials (cpu/material.h, material.cpp)
struct Lobe {
    int type = TPT_BXDF_LAMBERTIAN;
    float param = 0;
    float alpha2 = 0; // GGX
    Tex color;
    bool specular = false;
};
struct Mat {
    Lobe lobe[2];
    int combine = TPT_COMBINE_NONE;
    float weight = .5f;
    Tex emission, alpha;
    int bump = -1;
    bool specular = false;     // surface()->is_specular()
    bool alpha_opaque = true;  // AlphaTexture::empty()
    bool emissive = false;     // !emission_texture().empty()
};

struct Hit { // cpu/ray.h:31-45
    float t = INF;
    V3 normal{0, 0, 0}, shade_normal{0, 0, 0};
    bool into = true;
    V2 uv{0, 0};
    int mat = -1;
    int32_t prim = -1;
    bool is_hit() const { return t < INF; }
};

struct Ray {
    V3 org, dir;
    V3 at(float t) const { return org + dir * t; } // ray.h:17
};

// ---------------------------------------------------------------- primitives
enum { P_SPHERE = 0, P_RECT = 1, P_CIRCLE = 2, P_TRI = 3 };
struct Prim {
    int type;
    int mat;
    int32_t gid;
    V3 center;
    float radius;
    Rot rot;
    V2 dim;
    float dist;
    // triangle
    V3 a, ab, ac;
    int mesh, face;
};

struct AABB {
    V3 mn, mx;
    // AABB::intersect (cpu/bvh.cpp:8-27)
    float intersect(const Ray &r) const {
        float tmin = -INF, tmax = INF;
        for (int a = 0; a < 3; a++) {
            float inv = 1.f / r.dir[a];
            float t0 = (mn[a] - r.org[a]) * inv;
            float t1 = (mx[a] - r.org[a]) * inv;
            if (inv < 0.f) std::swap(t0, t1);
            tmin = t0 > tmin ? t0 : tmin;
            tmax = t1 < tmax ? t1 : tmax;
        }
        if (tmax < tmin || tmax < EPS) return INF;
        return tmin;
    }
    void extend(const AABB &o) {
        mn = cwise_min(mn, o.mn);
        mx = cwise_max(mx, o.mx);
    }
};

struct BVHNode {
    AABB box;
    int left = -1, right = -1;
    std::vector<int> objs; // prim indices
};
struct BVHTree { // cpu/bvh.h:55-86
    std::vector<BVHNode> nodes; // nodes[0] = root when non-empty
    bool empty() const { return nodes.empty(); }
};

struct MeshData {
    std::vector<V3> normals;
    std::vector<V2> texcoords;
    std::vector<int32_t> faces; // 9 per face
};

struct Light {
    int kind;   // 0 prim, 1 env, 2 whole mesh (ObjectLight of a TriangleMesh)
    int prim;   // prim index (kind 0), top-level object index (kind 2)
};
struct DeltaLight {
    int type;
    Rot rot;
    V3 center, color;
    float power, radius, angular_radius;
    V3 emission() const { return color * power; } // light.h:67,100
};

struct TopObj {
    int type;   // P_SPHERE/P_RECT/P_CIRCLE -> prim index; mesh -> shapes
    int prim;
    std::vector<BVHTree> shapes; // one BVHTree(faces, 4) per shape (cpu/object.cpp:606)
    std::vector<std::vector<int>> shape_faces;
};

struct Camera { // cpu/camera.cpp:6-21
    V3 center;
    Rot rot;
    int w, h;
    float fov, aperture, focus;
    V3 direct, up, right;
};

struct AlphaCtx {
    uint64_t key;
    int depth;
};

} // namespace

struct oracle_scene {
    std::vector<Img> images;
    std::vector<Mat> mats;
    std::vector<Prim> prims;
    std::vector<MeshData> meshes;
    std::vector<TopObj> top;
    std::vector<DeltaLight> delta;
    std::vector<Light> obj_lights;
    Camera cam;
    Tex bg;
    Rot bg_rot;
    bool bg_map = false;
    Dist2D env_dist;

    // ------------------------------------------------ textures
    V3 color_at(const Tex &t, V2 uv) const { // texture.cpp:16-23
        if (t.image < 0) return t.value;
        const Img &im = images[t.image];
        int x, y;
        denormalize(uv, im.w, im.h, x, y);
        return {im.texel(x, y, t.channel), im.texel(x, y, t.channel + 1), im.texel(x, y, t.channel + 2)};
    }
    float alpha_at(const Tex &t, V2 uv) const { // texture.cpp:25-32
        if (t.image < 0) return t.value.x;
        const Img &im = images[t.image];
        int x, y;
        denormalize(uv, im.w, im.h, x, y);
        return im.texel(x, y, t.channel);
    }
    // BumpTexture::bump_normal (texture.cpp:52-63)
    V3 bump_normal(int img, V2 uv, V3 n, V3 dpdu, V3 dpdv) const {
        const Img &im = images[img];
        int x, y;
        denormalize(uv, im.w, im.h, x, y);
        float dbdu = im.texel(x, y, 0), dbdv = im.texel(x, y, 1);
        V3 delta = dpdv * dbdv + dpdu * dbdu;
        delta = delta / (std::sqrt(std::sqrt(sqnorm(dpdu) * sqnorm(dpdv))) + EPS);
        return normalized(n - delta);
    }
    V3 background(V3 dir) const { // EnvTexture::color_at (texture.cpp:92-100)
        V2 uv{0, 0};
        if (bg_map) uv = dir2uv(mat_apply(bg_rot, dir));
        return color_at(bg, uv);
    }

    // ------------------------------------------------ primitive intersection
    template <bool SURF>
    void isect_sphere(const Prim &p, const Ray &ray, Hit &hit, const AlphaCtx *ctx) const { // object.cpp:21-64
        V3 op = p.center - ray.org;
        float b = dot(op, ray.dir);
        float det = square(b) - dot(op, op) + square(p.radius);
        if (det < 0) return;
        det = std::sqrt(det);
        float t;
        bool into;
        if (b - det >= EPS) {
            t = b - det;
            into = true;
        } else if (b + det >= EPS) {
            t = b + det;
            into = false;
        } else {
            t = INF;
            into = true;
        }
        if (t < hit.t) {
            if (SURF) {
                V3 pos = ray.at(t);
                V3 n = normalized(pos - p.center);
                float u = (std::atan2(n.y, n.x) + PI_F) * (.5f * INV_PI_F);
                float v = std::acos(n.z) * INV_PI_F;
                V2 uv{u, v};
                const Mat &m = mats[p.mat];
                if (ctx && !m.alpha_opaque && alpha_at(m.alpha, uv) < alpha_draw(ctx->key, ctx->depth, p.gid)) return;
                if (!into) n = -n;
                hit.t = t;
                hit.normal = n;
                hit.shade_normal = n;
                hit.into = into;
                hit.uv = uv;
                hit.mat = p.mat;
                hit.prim = p.gid;
            } else {
                hit.t = t;
                hit.prim = p.gid;
            }
        }
    }
    template <bool SURF>
    void isect_rect(const Prim &p, const Ray &ray, Hit &hit) const { // object.cpp:114-131
        const V3 &n = p.rot.w;
        float t = (p.dist - dot(ray.org, n)) / dot(ray.dir, n);
        if (EPS <= t && t < hit.t) {
            V3 cp = ray.at(t) - p.center;
            float u = dot(cp, p.rot.u) / p.dim.x + 0.5f;
            float v = dot(cp, p.rot.v) / p.dim.y + 0.5f;
            if (0 <= u && u <= 1 && 0 <= v && v <= 1) {
                hit.t = t;
                hit.prim = p.gid;
                if (SURF) {
                    V3 hn = same_side(n, ray.dir) ? -n : n;
                    hit.normal = hn;
                    hit.shade_normal = hn;
                    hit.into = true;
                    hit.uv = {u, v};
                    hit.mat = p.mat;
                }
            }
        }
    }
    template <bool SURF>
    void isect_circle(const Prim &p, const Ray &ray, Hit &hit) const { // object.cpp:163-180
        const V3 &n = p.rot.w;
        float t = (p.dist - dot(ray.org, n)) / dot(ray.dir, n);
        if (EPS <= t && t < hit.t) {
            V3 cp = ray.at(t) - p.center;
            if (sqnorm(cp) <= square(p.radius)) {
                hit.t = t;
                hit.prim = p.gid;
                if (SURF) {
                    V3 hn = same_side(n, ray.dir) ? -n : n;
                    cp = cp / p.radius;
                    hit.normal = hn;
                    hit.shade_normal = hn;
                    hit.into = true;
                    hit.uv = {dot(cp, p.rot.u), dot(cp, p.rot.v)};
                    hit.mat = p.mat;
                }
            }
        }
    }
    template <bool SURF>
    void isect_tri(const Prim &p, const Ray &ray, Hit &hit, const AlphaCtx *ctx) const { // object.cpp:250-336
        float t, u, v;
        {
            V3 pv = cross(ray.dir, p.ac);
            float inv_det = 1.f / dot(pv, p.ab);
            V3 ao = ray.org - p.a;
            u = dot(pv, ao) * inv_det;
            V3 q = cross(ao, p.ab);
            v = dot(q, ray.dir) * inv_det;
            if (!(0 <= u && 0 <= v && u + v <= 1)) return;
            t = dot(q, p.ac) * inv_det;
        }
        if (EPS <= t && t < hit.t) {
            if (SURF) {
                const MeshData &md = meshes[p.mesh];
                const int32_t *f = &md.faces[(size_t)p.face * 9];
                V2 uv;
                bool has_tc = f[2] >= 0;
                if (has_tc) {
                    V2 ta = md.texcoords[f[2]], tb = md.texcoords[f[5]], tc = md.texcoords[f[8]];
                    float w0 = 1 - u - v;
                    uv = {(w0 * ta.x + u * tb.x) + v * tc.x, (w0 * ta.y + u * tb.y) + v * tc.y};
                } else {
                    uv = {u, v};
                }
                const Mat &m = mats[p.mat];
                if (ctx && !m.alpha_opaque && alpha_at(m.alpha, uv) < alpha_draw(ctx->key, ctx->depth, p.gid)) return;
                V3 n;
                if (f[1] >= 0) {
                    V3 na = md.normals[f[1]], nb = md.normals[f[4]], nc = md.normals[f[7]];
                    float w0 = 1 - u - v;
                    n = normalized((w0 * na + u * nb) + v * nc);
                } else {
                    n = normalized(cross(p.ac, p.ab));
                }
                bool into = !same_side(n, ray.dir);
                if (!into) n = -n;
                V3 sn = n;
                if (m.bump >= 0) {
                    V2 dab, dac;
                    if (has_tc) {
                        V2 ta = md.texcoords[f[2]], tb = md.texcoords[f[5]], tc = md.texcoords[f[8]];
                        dab = {tb.x - ta.x, tb.y - ta.y};
                        dac = {tc.x - ta.x, tc.y - ta.y};
                    } else {
                        dab = {1, 0};
                        dac = {0, 1};
                    }
                    float inv_det = 1.f / (dab.x * dac.y - dac.x * dab.y);
                    if (std::isfinite(inv_det)) {
                        V3 dpdu = (dac.y * p.ab - dab.y * p.ac) * inv_det;
                        V3 dpdv = (-dac.x * p.ab + dab.x * p.ac) * inv_det;
                        sn = bump_normal(m.bump, uv, sn, dpdu, dpdv);
                    }
                }
                hit.t = t;
                hit.normal = n;
                hit.shade_normal = sn;
                hit.into = into;
                hit.uv = uv;
                hit.mat = p.mat;
                hit.prim = p.gid;
            } else {
                hit.t = t;
                hit.prim = p.gid;
            }
        }
    }
    template <bool SURF>
    void isect_prim(const Prim &p, const Ray &ray, Hit &hit, const AlphaCtx *ctx) const {
        switch (p.type) {
        case P_SPHERE: isect_sphere<SURF>(p, ray, hit, ctx); break;
        case P_RECT: isect_rect<SURF>(p, ray, hit); break;
        case P_CIRCLE: isect_circle<SURF>(p, ray, hit); break;
        default: isect_tri<SURF>(p, ray, hit, ctx); break;
        }
    }

    // BVHTree::intersect_impl / intersect_node (cpu/bvh.cpp:78-118)
    template <bool SURF>
    void bvh_node(const BVHTree &tree, int ni, const Ray &ray, Hit &hit, const AlphaCtx *ctx) const {
        const BVHNode &node = tree.nodes[ni];
        if (node.left < 0) {
            for (int o : node.objs) isect_prim<SURF>(prims[o], ray, hit, ctx);
            return;
        }
        float lt = tree.nodes[node.left].box.intersect(ray);
        float rt = tree.nodes[node.right].box.intersect(ray);
        if (lt < rt) {
            if (lt < hit.t) bvh_node<SURF>(tree, node.left, ray, hit, ctx);
            if (rt < hit.t) bvh_node<SURF>(tree, node.right, ray, hit, ctx);
        } else {
            if (rt < hit.t) bvh_node<SURF>(tree, node.right, ray, hit, ctx);
            if (lt < hit.t) bvh_node<SURF>(tree, node.left, ray, hit, ctx);
        }
    }
    template <bool SURF>
    void bvh_isect(const BVHTree &tree, const Ray &ray, Hit &hit, const AlphaCtx *ctx) const {
        if (tree.empty()) return;
        float t = tree.nodes[0].box.intersect(ray);
        if (t < hit.t) bvh_node<SURF>(tree, 0, ray, hit, ctx);
    }
    // Scene::intersect -> ObjectGroup::intersect_impl, linear top level (cpu/object.cpp:389-411)
    template <bool SURF>
    void intersect(const Ray &ray, Hit &hit, const AlphaCtx *ctx) const {
        for (const auto &o : top) {
            if (o.type == TPT_OBJ_MESH) {
                for (size_t s = 0; s < o.shapes.size(); s++) {
                    if (!o.shapes[s].empty()) {
                        bvh_isect<SURF>(o.shapes[s], ray, hit, ctx);
                    } else {
                        for (int pi : o.shape_faces[s]) isect_prim<SURF>(prims[pi], ray, hit, ctx);
                    }
                }
            } else {
                isect_prim<SURF>(prims[o.prim], ray, hit, ctx);
            }
        }
    }
    bool occluded(const Ray &ray, float tmax) const {
        Hit h;
        h.t = tmax;
        intersect<false>(ray, h, nullptr);
        return h.t < tmax;
    }

    // ------------------------------------------------ light sampling (cpu/object.cpp, light.cpp)
    float cos_theta_max(const Prim &p, V3 org) const { // object.cpp:95-104
        float d2 = sqnorm(p.center - org);
        float r2 = square(p.radius);
        if (d2 <= r2) return -1;
        return std::sqrt(1 - r2 / d2);
    }
    V3 prim_sample(const Prim &p, V3 org, RandEngine &rng) const {
        switch (p.type) {
        case P_SPHERE: { // object.cpp:66-75
            float ctm = cos_theta_max(p, org);
            float ct = rng.random(ctm, 1);
            float st = std::sqrt(1 - ct * ct);
            float phi = rng.random(2.f * PI_F);
            V3 local{st * std::cos(phi), st * std::sin(phi), ct};
            return rot_apply(from_direction(normalized(p.center - org)), local);
        }
        case P_RECT: { // object.cpp:133-138
            V3 uvec = ((rng.random() - .5f) * p.dim.x) * p.rot.u;
            V3 vvec = ((rng.random() - .5f) * p.dim.y) * p.rot.v;
            V3 target = (p.center + uvec) + vvec;
            return normalized(target - org);
        }
        case P_CIRCLE: { // object.cpp:188-193
            V2 uv = rng.rand_in_disk(p.radius);
            V3 target = (p.center + p.rot.u * uv.x) + p.rot.v * uv.y;
            return normalized(target - org);
        }
        default: { // object.cpp:226-235
            float sr1 = std::sqrt(rng.random());
            float r2 = rng.random();
            float u = sr1 * (1 - r2);
            float v = r2 * sr1;
            V3 target = (p.a + u * p.ab) + v * p.ac;
            return normalized(target - org);
        }
        }
    }
    static float plane_pdf(float dist, float area, V3 n, V3 dir) { // object.cpp:11-16
        float d2 = square(dist);
        float ca = std::fabs(dot(n, dir));
        float pdf = d2 / (ca * area);
        return (pdf < INF) ? pdf : INF;
    }
    float prim_pdf(const Prim &p, const Ray &ray) const {
        Hit h;
        switch (p.type) {
        case P_SPHERE: { // object.cpp:77-88
            isect_sphere<false>(p, ray, h, nullptr);
            if (!h.is_hit()) return 0;
            float ctm = cos_theta_max(p, ray.org);
            float sa = 2 * PI_F * (1 - ctm);
            return 1 / sa;
        }
        case P_RECT: // object.cpp:140-148
            isect_rect<true>(p, ray, h);
            if (!h.is_hit()) return 0;
            return plane_pdf(h.t, p.dim.x * p.dim.y, p.rot.w, ray.dir);
        case P_CIRCLE: // object.cpp:195-203
            isect_circle<false>(p, ray, h);
            if (!h.is_hit()) return 0;
            return plane_pdf(h.t, PI_F * square(p.radius), p.rot.w, ray.dir);
        default: { // object.cpp:237-248
            isect_tri<false>(p, ray, h, nullptr);
            if (!h.is_hit()) return 0;
            V3 fn = cross(p.ac, p.ab);
            float area = .5f * norm(fn);
            fn = normalized(fn);
            return plane_pdf(h.t, area, fn, ray.dir);
        }
        }
    }
    V3 env_sample(RandEngine &rng) const { // texture.cpp:79-84 (Distribution2D::sample(rng), distribution.h:41)
        float u = rng.random();
        float v = rng.random();
        V2 uv = env_dist.sample(u, v);
        return mat_apply_t(bg_rot, uv2dir(uv));
    }
    float env_pdf(V3 dir) const { // texture.cpp:86-90
        V2 uv = dir2uv(mat_apply(bg_rot, dir));
        return env_dist.pdf(uv) * (0.25f * INV_PI_F);
    }
    // ObjectLight of a mesh: ObjectGroup::sample / pdf (object.cpp:352-365) over the
    // mesh's shapes, each an ObjectGroup over its faces (shapes without faces left out)
    V3 mesh_sample(const TopObj &t, V3 pos, RandEngine &rng) const {
        std::vector<const std::vector<int> *> shapes;
        for (const auto &f : t.shape_faces)
            if (!f.empty()) shapes.push_back(&f);
        const std::vector<int> &faces = *shapes[rng.rand_int((int)shapes.size())];
        return prim_sample(prims[faces[rng.rand_int((int)faces.size())]], pos, rng);
    }
    float mesh_pdf(const TopObj &t, const Ray &r) const {
        float val = 0;
        int n = 0;
        for (const auto &faces : t.shape_faces) {
            if (faces.empty()) continue;
            float sv = 0;
            for (int pi : faces) sv += prim_pdf(prims[pi], r);
            val += sv / (float)faces.size();
            n++;
        }
        return val / (float)n;
    }
    V3 lights_sample(V3 pos, RandEngine &rng) const { // LightGroup::sample (light.cpp:35-39)
        const Light &l = obj_lights[rng.rand_int((int)obj_lights.size())];
        if (l.kind == 1) return env_sample(rng);
        if (l.kind == 2) return mesh_sample(top[l.prim], pos, rng);
        return prim_sample(prims[l.prim], pos, rng);
    }
    float lights_pdf(V3 pos, V3 dir) const { // LightGroup::pdf (light.cpp:41-48)
        float val = 0;
        Ray r{pos, dir};
        for (const Light &l : obj_lights)
            val += (l.kind == 1) ? env_pdf(dir) : (l.kind == 2) ? mesh_pdf(top[l.prim], r) : prim_pdf(prims[l.prim], r);
        return val / (float)obj_lights.size();
    }
    // DistantLight::sample / PointLight::sample (light.cpp:11-27); tmax = distance bound
    V3 delta_sample(const DeltaLight &l, V3 pos, RandEngine &rng, float &tmax) const {
        if (l.type == TPT_LIGHT_DISTANT) {
            tmax = INF;
            if (l.angular_radius < EPS) return l.rot.w;
            float du = rng.random(-l.angular_radius, l.angular_radius);
            float dv = rng.random(-l.angular_radius, l.angular_radius);
            return normalized(rot_apply(l.rot, V3{du, dv, 1.f}));
        }
        V3 c = l.center;
        if (l.radius >= EPS) c = c + rng.rand_on_sphere(l.radius);
        V3 d = c - pos;
        tmax = norm(d);
        return normalized(d);
    }

    // ------------------------------------------------ BxDFs (cpu/material.cpp)
    V3 lobe_f(const Lobe &b, const Ray &ray, const Hit &hit, V3 out) const {
        switch (b.type) {
        case TPT_BXDF_LAMBERTIAN: { // material.cpp:6-13
            float c = dot(hit.shade_normal, out);
            if (c > 0) return (c * INV_PI_F) * color_at(b.color, hit.uv);
            return {0, 0, 0};
        }
        case TPT_BXDF_METAL:
        case TPT_BXDF_DIELECTRIC: return color_at(b.color, hit.uv); // material.cpp:163,175-177
        default: { // Glossy::f (material.cpp:123-134)
            V3 l = -ray.dir;
            V3 v = out;
            V3 n = hit.shade_normal;
            if (!same_side(v, n)) return {0, 0, 0};
            V3 h = normalized(v + l);
            V3 col = color_at(b.color, hit.uv);
            float D = micro_D(b, n, h);
            float G = micro_G(b, l, n, v);
            float s = D * G / (4 * dot(n, l) * dot(n, v));
            float vh = 1.f - dot(v, h);
            float p5 = (vh * vh) * (vh * vh) * vh;
            V3 F{col.x + (1 - col.x) * p5, col.y + (1 - col.y) * p5, col.z + (1 - col.z) * p5};
            V3 d = F * s;
            return all_finite(d) ? d : V3{0, 0, 0};
        }
        }
    }
    static float micro_D(const Lobe &b, V3 n, V3 h) {
        if (b.type == TPT_BXDF_GLOSSY_BLINN) { // material.cpp:70-73
            float cb = dot(h, n);
            return (cb <= 0) ? 0 : (b.param + 2) * (.5f * INV_PI_F) * std::pow(cb, b.param + 1);
        }
        float cb = dot(n, h); // material.cpp:103-111
        if (cb <= 0) return 0;
        float c2 = cb * cb;
        float s2 = 1 - c2;
        return b.alpha2 * INV_PI_F * cb / square(b.alpha2 * c2 + s2);
    }
    static float micro_G(const Lobe &b, V3 l, V3 n, V3 v) {
        if (b.type == TPT_BXDF_GLOSSY_BLINN) { // material.cpp:37-56
            auto G1 = [&](V3 w) {
                float ct = dot(w, n);
                ct = std::max(std::min(ct, 1.f), 0.f);
                float st = std::sqrt(1 - ct * ct);
                float a = std::sqrt(.5f * b.param + 1) * ct / st;
                if (a < 1.6) {
                    float a2 = a * a;
                    return (3.535f * a + 2.181f * a2) / (1 + 2.276f * a + 2.577f * a2);
                }
                return 1.f;
            };
            return G1(l) * G1(v);
        }
        auto G1 = [&](V3 w) { // material.cpp:77-86
            float ct = dot(w, n);
            float tan2 = 1 / square(ct) - 1;
            return 2 / (1 + std::sqrt(1 + b.alpha2 * tan2));
        };
        return G1(l) * G1(v);
    }
    // microfacet normal: BlinnPhongModel::sample_h (material.cpp:58-66) and
    // TrowbridgeReitzModel::sample_h (:88-99)
    static V3 micro_sample_h(const Lobe &b, V3 n, RandEngine &rng) {
        float cb, sb, phi;
        if (b.type == TPT_BXDF_GLOSSY_BLINN) {
            cb = std::pow(rng.random(), 1 / (b.param + 2));
            sb = std::sqrt(1 - cb * cb);
            phi = rng.random(2.f * PI_F);
        } else {
            float u = rng.random();
            float c2 = (1 - u) / (1 - (1 - b.alpha2) * u);
            cb = std::sqrt(c2);
            sb = std::sqrt(1 - c2);
            phi = rng.random(2 * PI_F);
        }
        Rot r = from_direction(n);
        return rot_apply(r, V3{std::cos(phi) * sb, std::sin(phi) * sb, cb});
    }
    V3 lambert_sample(const Hit &hit, RandEngine &rng) const { // material.cpp:23-33
        float phi = rng.random(2.f * PI_F);
        float s2 = rng.random();
        float st = std::sqrt(s2);
        float ct = std::sqrt(1 - s2);
        Rot r = from_direction(hit.normal);
        return rot_apply(r, V3{std::cos(phi) * st, std::sin(phi) * st, ct});
    }
    V3 lobe_sample(const Lobe &b, const Ray &ray, const Hit &hit, RandEngine &rng) const {
        switch (b.type) {
        case TPT_BXDF_LAMBERTIAN: return lambert_sample(hit, rng);
        case TPT_BXDF_METAL: return reflect(ray.dir, hit.normal); // material.cpp:165-169
        case TPT_BXDF_DIELECTRIC: { // material.cpp:179-208
            V3 refl = reflect(ray.dir, hit.normal);
            const float nc = 1, ng = b.param;
            float eta = hit.into ? (nc / ng) : (ng / nc);
            float cosi = -dot(ray.dir, hit.normal);
            float cos2t = 1 - eta * eta * (1 - cosi * cosi);
            if (cos2t < 0) return refl;
            float cost = std::sqrt(cos2t);
            V3 refr = refract(ray.dir, hit.normal, eta, cosi, cost);
            float R0 = square(ng - nc) / square(ng + nc);
            float c = 1 - (hit.into ? cosi : cost);
            float Re = R0 + (1 - R0) * c * c * c * c * c;
            return (rng.random() < Re) ? refl : refr;
        }
        default: { // Glossy::sample (material.cpp:136-144)
            V3 h = micro_sample_h(b, hit.normal, rng);
            V3 out = reflect(ray.dir, h);
            if (!same_side(out, hit.normal)) return lambert_sample(hit, rng);
            return out;
        }
        }
    }
    float lobe_pdf(const Lobe &b, const Ray &ray, const Hit &hit, V3 out) const {
        if (b.type == TPT_BXDF_LAMBERTIAN) { // material.cpp:17-21
            float c = dot(out, hit.normal);
            return (c > 0) ? c * INV_PI_F : 0;
        }
        // Glossy::pdf (material.cpp:146-155); Metal/Dielectric pdf are never reached (specular)
        V3 l = -ray.dir;
        V3 n = hit.normal;
        if (!same_side(out, n)) return 0;
        V3 h = normalized(out + l);
        return micro_D(b, n, h) / (4 * dot(n, l) * dot(n, out));
    }
    // surface-level dispatch; sel = -1 is the combined (non-specular) Add/Mix BXDF
    V3 bxdf_f(const Mat &m, int sel, const Ray &ray, const Hit &hit, V3 out) const {
        if (sel >= 0) return lobe_f(m.lobe[sel], ray, hit, out);
        V3 f0 = lobe_f(m.lobe[0], ray, hit, out), f1 = lobe_f(m.lobe[1], ray, hit, out);
        if (m.combine == TPT_COMBINE_ADD) return f0 + f1;          // material.cpp:223-226
        return m.weight * f0 + (1 - m.weight) * f1;                // material.cpp:265-268
    }
    V3 bxdf_sample(const Mat &m, int sel, const Ray &ray, const Hit &hit, RandEngine &rng) const {
        if (sel >= 0) return lobe_sample(m.lobe[sel], ray, hit, rng);
        float thr = (m.combine == TPT_COMBINE_ADD) ? .5f : m.weight; // material.cpp:228-231, 270-273
        return (rng.random() < thr) ? lobe_sample(m.lobe[0], ray, hit, rng) : lobe_sample(m.lobe[1], ray, hit, rng);
    }
    float bxdf_pdf(const Mat &m, int sel, const Ray &ray, const Hit &hit, V3 out) const {
        if (sel >= 0) return lobe_pdf(m.lobe[sel], ray, hit, out);
        float p0 = lobe_pdf(m.lobe[0], ray, hit, out), p1 = lobe_pdf(m.lobe[1], ray, hit, out);
        if (m.combine == TPT_COMBINE_ADD) return .5f * (p0 + p1);   // material.cpp:233-236
        return m.weight * p0 + (1 - m.weight) * p1;                 // material.cpp:275-278
    }
    // sample_bxdf / pdf_bxdf (material.h:29-30, material.cpp:213-221, 244-263)
    int sample_bxdf(const Mat &m, RandEngine &rng, float &pdf_bxdf, bool &specular) const {
        if (m.combine == TPT_COMBINE_NONE) {
            pdf_bxdf = 1;
            specular = m.lobe[0].specular;
            return 0;
        }
        if (!m.specular) {
            pdf_bxdf = 1;
            specular = false;
            return -1;
        }
        int sel;
        if (m.combine == TPT_COMBINE_ADD) {
            sel = (rng.random() < .5f) ? 0 : 1;
            pdf_bxdf = 0.5f;
        } else {
            sel = (rng.random() < m.weight) ? 0 : 1;
            pdf_bxdf = sel == 0 ? m.weight : 1 - m.weight;
        }
        specular = m.lobe[sel].specular;
        return sel;
    }

    // ------------------------------------------------ camera (cpu/camera.cpp:23-34)
    Ray shoot(float x, float y, RandEngine &rng) const {
        V3 offset{0, 0, 0};
        if (cam.aperture >= EPS) {
            V2 o = rng.rand_in_disk(cam.aperture);
            offset = o.x * cam.rot.u + o.y * cam.rot.v;
        }
        V3 dir = (((x / (float)cam.w - .5f) * cam.right + (y / (float)cam.h - .5f) * cam.up) + cam.direct) - offset;
        return Ray{cam.center + offset, normalized(dir)};
    }

    // ------------------------------------------------ integrator (cpu/path_tracer.cpp:102-168)
    V3 radiance(Ray ray, RandEngine &rng, uint64_t key, const tpt_render_params &P) const {
        const float LIGHT_PROB = 1.f / 5; // cpu/path_tracer.cpp:104 (1 / reference MAX_DEPTH)
        int depth = 0;
        V3 decay{1, 1, 1}, color{0, 0, 0};
        while (true) {
            Hit hit;
            AlphaCtx ctx{key, depth};
            intersect<true>(ray, hit, &ctx);
            if (!hit.is_hit()) {
                color = color + cmul(decay, background(ray.dir));
                break;
            }
            const Mat &m = mats[hit.mat];
            color = color + cmul(decay, color_at(m.emission, hit.uv));
            if (depth >= P.max_depth) break;
            V3 hit_pos = ray.at(hit.t);
            V3 albedo, out;
            float pdf_bxdf;
            bool spec;
            int sel = sample_bxdf(m, rng, pdf_bxdf, spec);
            if (!spec) {
                if (!delta.empty()) {
                    const DeltaLight &dl = delta[rng.rand_int((int)delta.size())];
                    float tmax;
                    V3 to_light = delta_sample(dl, hit_pos, rng, tmax);
                    if (!P.shadow_bounded) tmax = INF;
                    if (!occluded(Ray{hit_pos, to_light}, tmax)) {
                        V3 la = bxdf_f(m, sel, ray, hit, to_light);
                        float pdf_light = 1.f / (float)delta.size();
                        color = color + cmul(cmul(decay, la), dl.emission()) / pdf_light;
                    }
                }
                float pdf_val;
                if (!obj_lights.empty()) {
                    if (rng.random() < LIGHT_PROB) {
                        out = lights_sample(hit_pos, rng);
                    } else {
                        out = bxdf_sample(m, sel, ray, hit, rng);
                    }
                    pdf_val = (1 - LIGHT_PROB) * bxdf_pdf(m, sel, ray, hit, out) +
                              LIGHT_PROB * lights_pdf(hit_pos, out);
                } else {
                    out = bxdf_sample(m, sel, ray, hit, rng);
                    pdf_val = bxdf_pdf(m, sel, ray, hit, out);
                }
                albedo = bxdf_f(m, sel, ray, hit, out) * (1.f / ((pdf_val + EPS) * pdf_bxdf));
            } else {
                out = bxdf_sample(m, sel, ray, hit, rng);
                albedo = bxdf_f(m, sel, ray, hit, out) * (1.f / pdf_bxdf);
            }
            decay = cmul(decay, albedo);
            if (is_zero(decay)) break;
            if (P.rr_depth >= 0 && depth >= P.rr_depth) { // Russian roulette (this build's extension)
                float mx = std::max(decay.x, std::max(decay.y, decay.z));
                float q = std::min(mx, 0.95f);
                if (rng.random() >= q) break;
                decay = decay / q;
            }
            ray = Ray{hit_pos, out};
            depth++;
        }
        return color;
    }

    // PixelSampler (cpu/path_tracer.cpp:8-58) + render loop body (:77-96)
    static int block_size(int spp) {
        const int m = 16;
        if (spp < 2 * 2 * m) return 1;
        if (spp < 3 * 3 * m) return 2;
        if (spp < 4 * 4 * m) return 3;
        return 4;
    }
    Ray sample_ray(int x, int y, int s, int b, int nsub, RandEngine &rng) const {
        int sub = s / nsub;
        int sx = sub % b, sy = sub / b;
        float w = 1.f / (float)b;
        float px0 = (float)x - .5f, py0 = (float)y - .5f;
        for (int i = 0; i < sx; i++) px0 += w;
        for (int i = 0; i < sy; i++) py0 += w;
        float px = px0 + rng.random(w);
        float py = py0 + rng.random(w);
        return shoot(px, py, rng);
    }
    void render_pixel(int x, int y, const tpt_render_params &P, float *out) const {
        int b = block_size(P.spp);
        int nsub = P.spp / (b * b);
        uint64_t pix = (uint64_t)y * (uint64_t)cam.w + (uint64_t)x;
        V3 color{0, 0, 0};
        double hdr[3] = {0, 0, 0};
        for (int sub = 0; sub < b * b; sub++) {
            V3 subc{0, 0, 0};
            double subd[3] = {0, 0, 0};
            for (int k = 0; k < nsub; k++) {
                int s = sub * nsub + k;
                uint64_t key = stream_seed(P.seed, pix, (uint64_t)s);
                RandEngine rng(key);
                Ray r = sample_ray(x, y, s, b, nsub, rng);
                V3 L = radiance(r, rng, key, P);
                if (P.accum_fp64) {
                    subd[0] += L.x;
                    subd[1] += L.y;
                    subd[2] += L.z;
                } else {
                    subc = subc + L;
                }
            }
            if (P.accum_fp64) {
                if (P.output == TPT_OUT_HDR_MEAN) {
                    for (int c = 0; c < 3; c++) hdr[c] += subd[c];
                } else {
                    V3 m{(float)(subd[0] / nsub), (float)(subd[1] / nsub), (float)(subd[2] / nsub)};
                    color = color + clip01(m);
                }
            } else {
                if (P.output == TPT_OUT_HDR_MEAN) {
                    color = color + subc;
                } else {
                    color = color + clip01(subc / (float)nsub);
                }
            }
        }
        V3 res;
        if (P.output == TPT_OUT_HDR_MEAN) {
            if (P.accum_fp64) {
                double n = (double)nsub * (double)(b * b);
                res = {(float)(hdr[0] / n), (float)(hdr[1] / n), (float)(hdr[2] / n)};
            } else {
                res = color / (float)(nsub * b * b);
            }
        } else {
            res = clip01(color / (float)(b * b));
        }
        float *o = out + ((size_t)y * cam.w + x) * 3;
        o[0] = res.x;
        o[1] = res.y;
        o[2] = res.z;
    }
};

// ---------------------------------------------------------------- construction
static Tex to_tex(const tpt_texture &t) {
    Tex r;
    r.image = t.image;
    r.channel = t.channel;
    r.value = {t.value[0], t.value[1], t.value[2]};
    return r;
}

// BVHTree(objects, node_size) + build_tree + get_bbox (cpu/bvh.cpp:41-76, 120-130)
static AABB prim_bbox(const Prim &p) {
    switch (p.type) {
    case P_SPHERE: { // object.cpp:90-93
        V3 o{p.radius, p.radius, p.radius};
        return {p.center - o, p.center + o};
    }
    case P_TRI: { // object.cpp:217-224
        V3 b = p.a + p.ab, c = p.a + p.ac;
        return {cwise_min(cwise_min(p.a, b), c), cwise_max(cwise_max(p.a, b), c)};
    }
    default: // rect / circle are never inside a mesh BVH here
        return {p.center, p.center};
    }
}

static void build_node(BVHTree &tree, int ni, std::vector<int> &objs, const std::vector<AABB> &boxes, int lo, int hi,
                       int height) {
    height--;
    AABB bb;
    if (lo == hi) {
        bb = {{0, 0, 0}, {0, 0, 0}};
    } else {
        bb = {{INF, INF, INF}, {-INF, -INF, -INF}};
        for (int i = lo; i < hi; i++) bb.extend(boxes[objs[i]]);
    }
    tree.nodes[ni].box = bb;
    if (height <= 0) {
        tree.nodes[ni].objs.assign(objs.begin() + lo, objs.begin() + hi);
        return;
    }
    V3 r = bb.mx - bb.mn;
    int axis = (r.x > r.y) ? ((r.x > r.z) ? 0 : 2) : ((r.y > r.z) ? 1 : 2);
    auto cmp = [&](int a, int b) { return boxes[a].mn[axis] < boxes[b].mn[axis]; };
    int mid = (lo + hi) / 2;
    std::nth_element(objs.begin() + lo, objs.begin() + mid, objs.begin() + hi, cmp);
    int l = (int)tree.nodes.size();
    tree.nodes.emplace_back();
    tree.nodes[ni].left = l;
    build_node(tree, l, objs, boxes, lo, mid, height);
    int rr = (int)tree.nodes.size();
    tree.nodes.emplace_back();
    tree.nodes[ni].right = rr;
    build_node(tree, rr, objs, boxes, mid, hi, height);
}

static BVHTree build_bvh(std::vector<int> objs, const std::vector<AABB> &boxes, size_t node_size) {
    BVHTree tree;
    if (objs.empty()) return tree;
    int height = 1;
    while ((node_size << (height - 1)) < objs.size()) height++;
    tree.nodes.emplace_back();
    build_node(tree, 0, objs, boxes, 0, (int)objs.size(), height);
    return tree;
}

static void set_lobe(Lobe &l, const tpt_bxdf &b) {
    l.type = b.type;
    l.param = b.param;
    l.color = to_tex(b.color);
    l.specular = (b.type == TPT_BXDF_METAL || b.type == TPT_BXDF_DIELECTRIC);
    if (b.type == TPT_BXDF_GLOSSY_GGX) l.alpha2 = 2 / (b.param + 2); // material.cpp:118
}

static oracle_scene *build_scene(const tpt_scene_desc *d) {
    auto s = std::make_unique<oracle_scene>();
    for (int i = 0; i < d->num_images; i++) {
        const tpt_image &im = d->images[i];
        Img g;
        g.w = im.width;
        g.h = im.height;
        g.c = im.channels;
        g.fmt = im.format;
        size_t n = (size_t)im.width * im.height * im.channels * (im.format == TPT_IMG_U8 ? 1 : 4);
        g.bytes.assign((const uint8_t *)im.data, (const uint8_t *)im.data + n);
        s->images.push_back(std::move(g));
    }
    for (int i = 0; i < d->num_materials; i++) {
        const tpt_material &m = d->materials[i];
        Mat g;
        set_lobe(g.lobe[0], m.lobe[0]);
        g.combine = m.combine;
        if (m.combine != TPT_COMBINE_NONE) {
            set_lobe(g.lobe[1], m.lobe[1]);
            if (m.combine == TPT_COMBINE_MIX && !(0 <= m.weight && m.weight <= 1))
                throw std::invalid_argument("[MixBXDF] invalid weight");
        }
        g.weight = m.weight;
        g.specular = g.lobe[0].specular || (m.combine != TPT_COMBINE_NONE && g.lobe[1].specular);
        g.emission = to_tex(m.emission);
        g.alpha = to_tex(m.alpha);
        g.bump = m.bump_image;
        g.alpha_opaque = (m.alpha.image < 0 && m.alpha.value[0] == 1.f);       // AlphaTexture::empty
        g.emissive = (m.emission.image >= 0) || !is_zero(g.emission.value);     // RGBTexture::empty
        s->mats.push_back(g);
    }
    int32_t gid = 0;
    for (int i = 0; i < d->num_objects; i++) {
        const tpt_object &o = d->objects[i];
        TopObj t;
        t.type = o.type;
        t.prim = -1;
        if (o.type == TPT_OBJ_MESH) {
            const tpt_mesh &m = d->meshes[o.mesh];
            int mi = (int)s->meshes.size();
            MeshData md;
            for (int k = 0; k < m.num_normals; k++)
                md.normals.push_back({m.normals[3 * k], m.normals[3 * k + 1], m.normals[3 * k + 2]});
            for (int k = 0; k < m.num_texcoords; k++) md.texcoords.push_back({m.texcoords[2 * k], m.texcoords[2 * k + 1]});
            md.faces.assign(m.faces, m.faces + (size_t)m.num_faces * 9);
            s->meshes.push_back(std::move(md));
            std::vector<AABB> boxes;
            int base = (int)s->prims.size();
            for (int f = 0; f < m.num_faces; f++) {
                const int32_t *fi = &m.faces[(size_t)f * 9];
                V3 va{m.vertices[3 * fi[0]], m.vertices[3 * fi[0] + 1], m.vertices[3 * fi[0] + 2]};
                V3 vb{m.vertices[3 * fi[3]], m.vertices[3 * fi[3] + 1], m.vertices[3 * fi[3] + 2]};
                V3 vc{m.vertices[3 * fi[6]], m.vertices[3 * fi[6] + 1], m.vertices[3 * fi[6] + 2]};
                Prim p{};
                p.type = P_TRI;
                p.mat = m.face_material[f];
                p.gid = gid + f;
                p.a = va;
                p.ab = vb - va; // TriangleMeshFace ctor (object.cpp:205-207)
                p.ac = vc - va;
                p.mesh = mi;
                p.face = f;
                s->prims.push_back(p);
            }
            for (int f = 0; f < m.num_faces; f++) boxes.push_back(prim_bbox(s->prims[base + f]));
            // boxes indexed by prim index: pad so boxes[prim] works
            std::vector<AABB> all(s->prims.size());
            for (int f = 0; f < m.num_faces; f++) all[base + f] = boxes[f];
            for (int sh = 0; sh < m.num_shapes; sh++) {
                std::vector<int> objs;
                for (int f = m.shape_offsets[sh]; f < m.shape_offsets[sh + 1]; f++) objs.push_back(base + f);
                t.shape_faces.push_back(objs);
                t.shapes.push_back(build_bvh(objs, all, 4));
            }
            gid += m.num_faces;
        } else {
            Prim p{};
            p.mat = o.material;
            p.gid = gid++;
            p.center = {o.center[0], o.center[1], o.center[2]};
            p.radius = o.radius;
            p.rot = rot_from9(o.rot);
            p.dim = {o.dim[0], o.dim[1]};
            if (o.type == TPT_OBJ_SPHERE) {
                p.type = P_SPHERE;
            } else if (o.type == TPT_OBJ_RECTANGLE) {
                p.type = P_RECT;
                p.dist = dot(p.rot.w, p.center); // Rectangle::calc_dist (object.h:116)
            } else if (o.type == TPT_OBJ_CIRCLE) {
                p.type = P_CIRCLE;
                p.dist = dot(p.rot.w, p.center); // object.cpp:157
            } else {
                throw std::invalid_argument("unknown object type");
            }
            t.prim = (int)s->prims.size();
            s->prims.push_back(p);
        }
        s->top.push_back(std::move(t));
    }
    // object lights: DFS with an explicit stack (cpu/scene.cpp:9-19)
    {
        struct Node {
            int kind; // 0 root, 1 top object, 2 mesh shape, 3 face/prim
            int a, b;
        };
        std::vector<Node> stk{{0, 0, 0}};
        while (!stk.empty()) {
            Node n = stk.back();
            stk.pop_back();
            if (n.kind == 0) {
                for (int i = 0; i < (int)s->top.size(); i++) stk.push_back({1, i, 0});
            } else if (n.kind == 1) {
                const TopObj &t = s->top[n.a];
                if (t.type == TPT_OBJ_MESH) {
                    for (int sh = 0; sh < (int)t.shapes.size(); sh++) stk.push_back({2, n.a, sh});
                } else if (s->mats[s->prims[t.prim].mat].emissive) {
                    s->obj_lights.push_back({0, t.prim});
                }
            } else if (n.kind == 2) {
                for (int pi : s->top[n.a].shape_faces[n.b]) stk.push_back({3, pi, 0});
            } else {
                if (s->mats[s->prims[n.a].mat].emissive) s->obj_lights.push_back({0, n.a});
            }
        }
    }
    // non-delta lights passed to Scene() join the mixture after the emissive objects (scene.cpp:21-27)
    for (int i = 0; i < d->num_object_lights; i++) {
        const tpt_object_light &ol = d->object_lights[i];
        if (ol.kind == TPT_OBJECT_LIGHT_ENV) {
            s->obj_lights.push_back({1, -1});
            continue;
        }
        if (ol.kind != TPT_OBJECT_LIGHT_OBJECT || ol.object < 0 || ol.object >= (int)s->top.size())
            throw std::invalid_argument("invalid object light");
        const TopObj &t = s->top[ol.object];
        if (t.type == TPT_OBJ_MESH) s->obj_lights.push_back({2, ol.object});
        else s->obj_lights.push_back({0, t.prim});
    }
    for (int i = 0; i < d->num_delta_lights; i++) {
        const tpt_delta_light &l = d->delta_lights[i];
        DeltaLight g;
        g.type = l.type;
        g.rot = rot_from9(l.rot);
        g.center = {l.center[0], l.center[1], l.center[2]};
        g.color = {l.color[0], l.color[1], l.color[2]};
        g.power = l.power;
        g.radius = l.radius;
        if (l.type == TPT_LIGHT_DISTANT) {
            if (!(0 <= l.radius && l.radius <= PI_F)) throw std::invalid_argument("angle out of range");
            g.angular_radius = l.radius / 2; // DistantLight ctor (light.cpp:6-9)
        }
        s->delta.push_back(g);
    }
    // background / EnvLight (scene.cpp:25-30)
    s->bg = to_tex(d->background.rgb);
    s->bg_rot = rot_from9(d->background.rot);
    s->bg_map = d->background.rgb.image >= 0;
    if (s->bg_map) {
        if (d->background.importance) { // the CDF texture.cpp:68-76 intended to build
            const Img &im = s->images[s->bg.image];
            std::vector<std::vector<float>> w(im.h, std::vector<float>(im.w));
            for (int y = 0; y < im.h; y++)
                for (int x = 0; x < im.w; x++) {
                    float r = im.texel(x, y, s->bg.channel), g = im.texel(x, y, s->bg.channel + 1),
                          b = im.texel(x, y, s->bg.channel + 2);
                    w[y][x] = std::max(r, std::max(g, b));
                }
            s->env_dist = Dist2D(w);
        }
        s->obj_lights.push_back({1, -1});
    }
    // camera (camera.cpp:6-21)
    const tpt_camera &c = d->camera;
    Camera &cam = s->cam;
    cam.center = {c.center[0], c.center[1], c.center[2]};
    cam.rot = rot_from9(c.rot);
    cam.w = c.width;
    cam.h = c.height;
    cam.fov = c.fov;
    cam.aperture = c.aperture;
    cam.focus = c.focus_dist;
    float vw, vh;
    if (c.width > c.height) {
        vw = 2.f * std::tan(c.fov * .5f);
        vh = vw * (float)c.height / (float)c.width;
    } else {
        vh = 2.f * std::tan(c.fov * .5f);
        vw = vh * (float)c.width / (float)c.height;
    }
    cam.direct = cam.focus * -cam.rot.w;
    cam.right = (cam.focus * vw) * cam.rot.u;
    cam.up = (cam.focus * vh) * cam.rot.v;
    return s.release();
}

static int omp_threads(int threads) {
#ifdef _OPENMP
    return threads > 0 ? threads : omp_get_max_threads();
#else
    (void)threads;
    return 1;
#endif
}

extern "C" {

const char *oracle_last_error(void) { return g_last_error.c_str(); }

int oracle_rng_run(uint64_t seed, const int32_t *ops, int32_t n, double *out) {
    RandEngine r(seed);
    for (int i = 0; i < n; i++) out[i] = ops[i] == 0 ? (double)r.random() : (double)r.rand_int();
    return 0;
}
void oracle_rng_state(uint64_t seed, uint16_t *st) {
    RandEngine r(seed);
    st[0] = r.s[0];
    st[1] = r.s[1];
    st[2] = r.s[2];
}
uint64_t oracle_stream_seed(uint64_t base, uint64_t pixel, uint64_t sample) { return stream_seed(base, pixel, sample); }

oracle_scene *oracle_scene_create(const tpt_scene_desc *desc) {
    try {
        return build_scene(desc);
    } catch (const std::exception &e) {
        g_last_error = e.what();
        return nullptr;
    }
}
void oracle_scene_destroy(oracle_scene *s) { delete s; }

int oracle_object_lights(const oracle_scene *s, int32_t *kinds, int32_t *ids, int32_t cap) {
    int n = (int)s->obj_lights.size();
    for (int i = 0; i < n && i < cap; i++) {
        kinds[i] = s->obj_lights[i].kind;
        const Light &l = s->obj_lights[i];
        ids[i] = l.kind == 0 ? s->prims[l.prim].gid : l.kind == 2 ? l.prim : -1;
    }
    return n;
}

int oracle_trace_closest(const oracle_scene *s, const float *rays, uint64_t n, int32_t *prim, float *t, int mode,
                         int threads) {
    int nt = omp_threads(threads);
#pragma omp parallel for schedule(dynamic, 4096) num_threads(nt)
    for (int64_t i = 0; i < (int64_t)n; i++) {
        const float *r = rays + 8 * i;
        Ray ray{{r[0], r[1], r[2]}, {r[4], r[5], r[6]}};
        Hit h;
        h.t = r[7] < INF ? r[7] : INF;
        if (mode == 0) {
            s->intersect<false>(ray, h, nullptr);
        } else {
            for (const Prim &p : s->prims) {
                Hit c;
                c.t = INF;
                s->isect_prim<false>(p, ray, c, nullptr);
                if (c.t < INF && (c.t < h.t || (c.t == h.t && c.prim < h.prim))) h = c;
            }
        }
        prim[i] = h.t < INF ? h.prim : -1;
        t[i] = h.t < INF ? h.t : INF;
    }
    return 0;
}

int oracle_trace_occluded(const oracle_scene *s, const float *rays, uint64_t n, uint8_t *occ, int threads) {
    int nt = omp_threads(threads);
#pragma omp parallel for schedule(dynamic, 4096) num_threads(nt)
    for (int64_t i = 0; i < (int64_t)n; i++) {
        const float *r = rays + 8 * i;
        Ray ray{{r[0], r[1], r[2]}, {r[4], r[5], r[6]}};
        occ[i] = s->occluded(ray, r[7] < INF ? r[7] : INF) ? 1 : 0;
    }
    return 0;
}

int32_t oracle_intersect_surface(const oracle_scene *s, const float *r, uint64_t key, int32_t depth, float *o) {
    Ray ray{{r[0], r[1], r[2]}, {r[4], r[5], r[6]}};
    Hit h;
    AlphaCtx ctx{key, depth};
    s->intersect<true>(ray, h, &ctx);
    o[0] = h.t;
    o[1] = h.normal.x;
    o[2] = h.normal.y;
    o[3] = h.normal.z;
    o[4] = h.shade_normal.x;
    o[5] = h.shade_normal.y;
    o[6] = h.shade_normal.z;
    o[7] = h.into ? 1.f : 0.f;
    o[8] = h.uv.x;
    o[9] = h.uv.y;
    return h.is_hit() ? h.prim : -1;
}

int oracle_intersect_surface_batch(const oracle_scene *s, const float *rays, uint64_t n, int32_t *prim, float *out10,
                                   int threads) {
    int nt = omp_threads(threads);
#pragma omp parallel for schedule(dynamic, 4096) num_threads(nt)
    for (int64_t i = 0; i < (int64_t)n; i++) prim[i] = oracle_intersect_surface(s, rays + 8 * i, 0, 0, out10 + 10 * i);
    return 0;
}

int oracle_render(const oracle_scene *s, const tpt_render_params *p, float *out, int32_t y0, int32_t y1, int threads) {
    int nt = omp_threads(threads);
    int W = s->cam.w;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nt)
    for (int y = y0; y < y1; y++)
        for (int x = 0; x < W; x++) s->render_pixel(x, y, *p, out);
    return 0;
}

int oracle_radiance_sample(const oracle_scene *s, const tpt_render_params *p, int32_t x, int32_t y, int32_t sample,
                           float *out3) {
    int b = oracle_scene::block_size(p->spp);
    int nsub = p->spp / (b * b);
    uint64_t pix = (uint64_t)y * (uint64_t)s->cam.w + (uint64_t)x;
    uint64_t key = stream_seed(p->seed, pix, (uint64_t)sample);
    RandEngine rng(key);
    Ray r = s->sample_ray(x, y, sample, b, nsub, rng);
    V3 L = s->radiance(r, rng, key, *p);
    out3[0] = L.x;
    out3[1] = L.y;
    out3[2] = L.z;
    return 0;
}

int oracle_bvh_octant_golden(int leaf_size, float *boxes, int32_t *leaf_obj, int32_t *is_leaf) {
    // bvh_test.cpp:9-23: 8 triangles, one per octant, objects t1..t8 = ids 0..7
    const float x1 = 6, y1 = 5, z1 = 4;
    V3 px{x1, 0, 0}, py{0, y1, 0}, pz{0, 0, z1};
    V3 tri[8][3] = {{px, py, pz},    {-px, py, pz},    {-px, -py, pz},    {px, -py, pz},
                    {px, py, -pz},   {-px, py, -pz},   {-px, -py, -pz},   {px, -py, -pz}};
    std::vector<Prim> ps(8);
    std::vector<AABB> bx(8);
    for (int i = 0; i < 8; i++) {
        ps[i] = Prim{};
        ps[i].type = P_TRI;
        ps[i].a = tri[i][0];
        ps[i].ab = tri[i][1] - tri[i][0];
        ps[i].ac = tri[i][2] - tri[i][0];
        bx[i] = prim_bbox(ps[i]);
    }
    BVHTree t = build_bvh({0, 1, 2, 3, 4, 5, 6, 7}, bx, (size_t)leaf_size);
    // heap order: index 1 = root, children 2i, 2i+1
    std::vector<int> heap(16, -1);
    heap[1] = 0;
    for (int i = 1; i < 8; i++) {
        if (heap[i] < 0) continue;
        const BVHNode &nd = t.nodes[heap[i]];
        if (nd.left >= 0) {
            heap[2 * i] = nd.left;
            heap[2 * i + 1] = nd.right;
        }
    }
    for (int i = 1; i < 16; i++) {
        int ni = heap[i];
        float *b = boxes + 6 * (i - 1);
        if (ni < 0) {
            for (int k = 0; k < 6; k++) b[k] = NAN;
            leaf_obj[i - 1] = -2;
            is_leaf[i - 1] = -1;
            continue;
        }
        const BVHNode &nd = t.nodes[ni];
        b[0] = nd.box.mn.x;
        b[1] = nd.box.mn.y;
        b[2] = nd.box.mn.z;
        b[3] = nd.box.mx.x;
        b[4] = nd.box.mx.y;
        b[5] = nd.box.mx.z;
        is_leaf[i - 1] = nd.left < 0;
        leaf_obj[i - 1] = (nd.left < 0 && nd.objs.size() == 1) ? nd.objs[0] : -1;
    }
    return (int)t.nodes.size();
}

int oracle_distribution1d(const float *w, int32_t n, float u, float *x, float *pdf) {
    try {
        Dist1D d(std::vector<float>(w, w + n));
        *x = d.sample(u);
        *pdf = d.pdf(*x);
        return 0;
    } catch (const std::exception &e) {
        g_last_error = e.what();
        return -1;
    }
}

int oracle_distribution2d(const float *w, int32_t rows, int32_t cols, float u, float v, float *xy, float *pdf) {
    try {
        std::vector<std::vector<float>> ww(rows, std::vector<float>(cols));
        for (int r = 0; r < rows; r++)
            for (int c = 0; c < cols; c++) ww[r][c] = w[r * cols + c];
        Dist2D d(ww);
        V2 s = d.sample(u, v);
        xy[0] = s.x;
        xy[1] = s.y;
        *pdf = d.pdf(s);
        return 0;
    } catch (const std::exception &e) {
        g_last_error = e.what();
        return -1;
    }
}

void oracle_env_uv2dir(float u, float v, float *d) {
    V3 r = uv2dir({u, v});
    d[0] = r.x;
    d[1] = r.y;
    d[2] = r.z;
}
void oracle_env_dir2uv(const float *d, float *uv) {
    V2 r = dir2uv({d[0], d[1], d[2]});
    uv[0] = r.x;
    uv[1] = r.y;
}

// cv::Sobel(ksize 3, dx|dy, scale, BORDER_REPLICATE) as used by BumpTexture (texture.cpp:34-44)
void oracle_sobel(const float *hm, int32_t w, int32_t h, float scale, float *grad) {
    scale /= 8;
    auto at = [&](int x, int y) {
        x = std::min(std::max(x, 0), w - 1);
        y = std::min(std::max(y, 0), h - 1);
        return hm[(size_t)y * w + x];
    };
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            float gx = (at(x + 1, y - 1) - at(x - 1, y - 1)) + 2.f * (at(x + 1, y) - at(x - 1, y)) +
                       (at(x + 1, y + 1) - at(x - 1, y + 1));
            float gy = (at(x - 1, y + 1) - at(x - 1, y - 1)) + 2.f * (at(x, y + 1) - at(x, y - 1)) +
                       (at(x + 1, y + 1) - at(x + 1, y - 1));
            grad[((size_t)y * w + x) * 2] = gx * scale;
            grad[((size_t)y * w + x) * 2 + 1] = gy * scale;
        }
}

void oracle_rotation_from_direction(const float *w3, float *m9) {
    Rot r = from_direction({w3[0], w3[1], w3[2]});
    float v[9] = {r.u.x, r.u.y, r.u.z, r.v.x, r.v.y, r.v.z, r.w.x, r.w.y, r.w.z};
    std::memcpy(m9, v, sizeof(v));
}

// Rotation::from_euler (cpu/mapping.cpp:32-48): R = AngleAxis(z) * AngleAxis(y) * AngleAxis(x)
int oracle_rotation_from_euler(const int32_t *axes, const float *angles, int32_t degrees, float *m9) {
    for (int i = 0; i < 3; i++)
        if (axes[i] < 0 || axes[i] > 2) {
            g_last_error = "expect all axes to be between 0 and 2";
            return -1;
        }
    double a[3];
    for (int i = 0; i < 3; i++) a[i] = degrees ? (double)(angles[i] * (PI_F / 180.f)) : (double)angles[i];
    auto axis_rot = [](int ax, double ang, double R[3][3]) {
        double c = std::cos(ang), s = std::sin(ang);
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) R[i][j] = (i == j) ? 1 : 0;
        int i1 = (ax + 1) % 3, i2 = (ax + 2) % 3;
        R[i1][i1] = c;
        R[i1][i2] = -s;
        R[i2][i1] = s;
        R[i2][i2] = c;
    };
    double Rz[3][3], Ry[3][3], Rx[3][3], T[3][3], R[3][3];
    axis_rot(axes[2], a[2], Rz);
    axis_rot(axes[1], a[1], Ry);
    axis_rot(axes[0], a[0], Rx);
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            T[i][j] = 0;
            for (int k = 0; k < 3; k++) T[i][j] += Rz[i][k] * Ry[k][j];
        }
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            R[i][j] = 0;
            for (int k = 0; k < 3; k++) R[i][j] += T[i][k] * Rx[k][j];
        }
    for (int col = 0; col < 3; col++)
        for (int row = 0; row < 3; row++) m9[col * 3 + row] = (float)R[row][col];
    return 0;
}

int oracle_delta_light_sample(const oracle_scene *s, int32_t light, const float *pos3, uint64_t seed, float *dir3) {
    if (light < 0 || light >= (int)s->delta.size()) return -1;
    RandEngine rng(seed);
    float tmax;
    V3 d = s->delta_sample(s->delta[light], {pos3[0], pos3[1], pos3[2]}, rng, tmax);
    dir3[0] = d.x;
    dir3[1] = d.y;
    dir3[2] = d.z;
    return 0;
}

int oracle_light_sample(const oracle_scene *s, int32_t li, const float *pos3, uint64_t seed, float *dir3, float *pdf) {
    if (li < 0 || li >= (int)s->obj_lights.size()) return -1;
    RandEngine rng(seed);
    V3 pos{pos3[0], pos3[1], pos3[2]};
    const Light &l = s->obj_lights[li];
    V3 d = (l.kind == 1)   ? s->env_sample(rng)
           : (l.kind == 2) ? s->mesh_sample(s->top[l.prim], pos, rng)
                           : s->prim_sample(s->prims[l.prim], pos, rng);
    dir3[0] = d.x;
    dir3[1] = d.y;
    dir3[2] = d.z;
    *pdf = (l.kind == 1)   ? s->env_pdf(d)
           : (l.kind == 2) ? s->mesh_pdf(s->top[l.prim], Ray{pos, d})
                           : s->prim_pdf(s->prims[l.prim], Ray{pos, d});
    return 0;
}

static Hit synth_hit(const float *n3, int into) {
    Hit h;
    h.t = 1;
    h.normal = {n3[0], n3[1], n3[2]};
    h.shade_normal = h.normal;
    h.into = into != 0;
    h.uv = {0.25f, 0.75f};
    return h;
}

int oracle_bxdf_eval(const oracle_scene *s, int32_t material, const float *rd, const float *n3, int32_t into,
                     const float *od, float *f3, float *pdf) {
    const Mat &m = s->mats[material];
    Ray ray{{0, 0, 0}, {rd[0], rd[1], rd[2]}};
    Hit h = synth_hit(n3, into);
    V3 out{od[0], od[1], od[2]};
    int sel = (m.combine == TPT_COMBINE_NONE) ? 0 : -1;
    V3 f = s->bxdf_f(m, sel, ray, h, out);
    f3[0] = f.x;
    f3[1] = f.y;
    f3[2] = f.z;
    *pdf = m.specular ? 0.f : s->bxdf_pdf(m, sel, ray, h, out);
    return 0;
}

int oracle_bxdf_sample(const oracle_scene *s, int32_t material, const float *rd, const float *n3, int32_t into,
                       uint64_t seed, float *od) {
    const Mat &m = s->mats[material];
    Ray ray{{0, 0, 0}, {rd[0], rd[1], rd[2]}};
    Hit h = synth_hit(n3, into);
    RandEngine rng(seed);
    float pb;
    bool spec;
    int sel = s->sample_bxdf(m, rng, pb, spec);
    V3 o = s->bxdf_sample(m, sel, ray, h, rng);
    od[0] = o.x;
    od[1] = o.y;
    od[2] = o.z;
    return 0;
}

static Lobe glossy_lobe(int32_t type, float shininess) {
    Lobe b;
    b.type = type;
    b.param = shininess;
    if (type == TPT_BXDF_GLOSSY_GGX) b.alpha2 = 2 / (shininess + 2); // material.cpp:118
    return b;
}

int oracle_micro_terms(int32_t type, float shininess, const float *n3, const float *l3, const float *v3,
                       const float *h3, float *D, float *G) {
    if (type != TPT_BXDF_GLOSSY_BLINN && type != TPT_BXDF_GLOSSY_GGX) return -1;
    Lobe b = glossy_lobe(type, shininess);
    V3 n{n3[0], n3[1], n3[2]}, l{l3[0], l3[1], l3[2]}, v{v3[0], v3[1], v3[2]}, h{h3[0], h3[1], h3[2]};
    *D = oracle_scene::micro_D(b, n, h);
    *G = oracle_scene::micro_G(b, l, n, v);
    return 0;
}

int oracle_micro_sample_h(int32_t type, float shininess, const float *n3, uint64_t seed, int32_t count, float *h3) {
    if (type != TPT_BXDF_GLOSSY_BLINN && type != TPT_BXDF_GLOSSY_GGX) return -1;
    Lobe b = glossy_lobe(type, shininess);
    V3 n{n3[0], n3[1], n3[2]};
    RandEngine rng(seed);
    for (int32_t i = 0; i < count; i++) {
        V3 h = oracle_scene::micro_sample_h(b, n, rng);
        h3[3 * i] = h.x;
        h3[3 * i + 1] = h.y;
        h3[3 * i + 2] = h.z;
    }
    return 0;
}

int oracle_camera_ray(const oracle_scene *s, float x, float y, uint64_t seed, float *org3, float *dir3) {
    RandEngine rng(seed);
    Ray r = s->shoot(x, y, rng);
    org3[0] = r.org.x;
    org3[1] = r.org.y;
    org3[2] = r.org.z;
    dir3[0] = r.dir.x;
    dir3[1] = r.dir.y;
    dir3[2] = r.dir.z;
    return 0;
}

} // extern "C"
