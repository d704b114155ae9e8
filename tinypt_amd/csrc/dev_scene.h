// dev_scene.h — device-side scene access: textures, primitive tests, light
// sampling and BxDFs of the HIP path tracer.  Each routine restates the
// reference CPU behaviour it names (li-plus/tinypt src/tinypt/cpu/*).
#pragma once

#include "dev_math.h"

namespace tpt {

// Read-only scene records at a wave-uniform address (light lists, analytic
// primitives, batch scene tables) are read through the constant address
// space: those loads are invariant, so they become scalar loads into SGPRs,
// like a kernel-argument record.  Through a generic pointer they would be
// vector loads, one per lane, repeated at every use.  With a divergent address
// the compiler falls back to vector loads, so the result is the same either way.
template <class T> __device__ __forceinline__ T load_invariant(const T *p) {
    static_assert(sizeof(T) % 4 == 0 && alignof(T) >= 4, "word-aligned records only");
    typedef const __attribute__((address_space(4))) uint32_t const_u32;
    const_u32 *src = (const_u32 *)p;
    T v;
    uint32_t *dst = reinterpret_cast<uint32_t *>(&v);
#pragma unroll
    for (int k = 0; k < (int)(sizeof(T) / 4); k++) dst[k] = src[k];
    return v;
}

// ---------------------------------------------------------------- textures (cpu/texture.cpp)
__device__ __forceinline__ void denorm_coords(V2 uv, int w, int h, int &x, int &y) { // texture.cpp:6-14
    float iu = uv.x - floorf(uv.x);
    float iv = uv.y - floorf(uv.y);
    x = (int)(iu * ((float)w - kEPS));
    y = (int)(iv * ((float)h - kEPS));
}
__device__ __forceinline__ float texel(const DScene &S, const DImage &im, int x, int y, int ch) {
    uint64_t idx = ((uint64_t)y * (uint64_t)im.w + (uint64_t)x) * (uint64_t)im.c + (uint64_t)ch;
    if (im.fmt == 1) return (float)S.texels[im.offset + idx] / 255.f;
    return reinterpret_cast<const float *>(S.texels + im.offset)[idx];
}
__device__ __forceinline__ V3 color_at(const DScene &S, const DTex &t, V2 uv) { // texture.cpp:16-23
    if (t.image < 0) return V3{t.v[0], t.v[1], t.v[2]};
    const DImage im = S.images[t.image];
    int x, y;
    denorm_coords(uv, im.w, im.h, x, y);
    return V3{texel(S, im, x, y, t.channel), texel(S, im, x, y, t.channel + 1), texel(S, im, x, y, t.channel + 2)};
}
__device__ __forceinline__ float alpha_at(const DScene &S, const DTex &t, V2 uv) { // texture.cpp:25-32
    if (t.image < 0) return t.v[0];
    const DImage im = S.images[t.image];
    int x, y;
    denorm_coords(uv, im.w, im.h, x, y);
    return texel(S, im, x, y, t.channel);
}
__device__ __forceinline__ V3 bump_normal(const DScene &S, int img, V2 uv, V3 n, V3 dpdu, V3 dpdv) { // :52-63
    const DImage im = S.images[img];
    int x, y;
    denorm_coords(uv, im.w, im.h, x, y);
    float dbdu = texel(S, im, x, y, 0), dbdv = texel(S, im, x, y, 1);
    V3 delta = add(scr(dpdv, dbdv), scr(dpdu, dbdu));
    delta = dvs(delta, __builtin_sqrtf(__builtin_sqrtf(sqn(dpdu) * sqn(dpdv))) + kEPS);
    return nrm(sub(n, delta));
}
__device__ __forceinline__ V2 dir2uv(V3 d) { // EnvTexture::dir2uv, texture.cpp:113-121
    float u = 1.f - (atan2f(d.y, d.x) + kPI) * (.5f * kINV_PI);
    float v = 1.f - acosf(fminf(fmaxf(d.z, -1.f), 1.f)) * kINV_PI;
    return {u, v};
}
__device__ __forceinline__ V3 uv2dir(V2 uv) { // texture.cpp:102-111
    float phi = 2.f * kPI * (1.f - uv.x) - kPI;
    float theta = kPI * (1.f - uv.y);
    float ct = cosf(theta);
    float st = __builtin_sqrtf(1.f - ct * ct);
    return {cosf(phi) * st, sinf(phi) * st, ct};
}
__device__ __forceinline__ V3 background(const DScene &S, V3 dir) { // EnvTexture::color_at
    V2 uv{0.f, 0.f};
    if (S.bg_map) uv = dir2uv(mat_apply(S.bg_rot, dir));
    return color_at(S, S.bg, uv);
}

// piecewise-constant CDF (cpu/distribution.cpp:32-44); cdf has count+1 entries.
// `guide` (G + 1 entries, G a power of two) narrows the upper_bound over
// [0, count] to [guide[k], guide[k + 1]] for k = floor(u * G): every index below
// guide[k] has cdf <= k / G <= u, and cdf[guide[k + 1]] > (k + 1) / G > u, so the
// narrowed search returns the same index in one or two steps instead of
// log2(count) dependent loads.
__device__ __forceinline__ float dist_sample(const float *cdf, int count, const int32_t *guide, int G, float u) {
    int k = (int)(u * (float)G); // exact: G is a power of two, u in [0, 1]
    k = k < G ? k : G - 1;
    int lo = guide[k], hi = guide[k + 1];
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (u < cdf[mid]) hi = mid; else lo = mid + 1;
    }
    int i = lo < count ? lo : count;
    return ((float)i - 1.f + (u - cdf[i - 1]) / (cdf[i] - cdf[i - 1])) / (float)count;
}
__device__ __forceinline__ int dist_anchor(int count, float x) { return (int)(x * ((float)count - kEPS)); }
__device__ __forceinline__ float dist_pdf(const float *cdf, int count, float x) {
    int i = dist_anchor(count, x);
    return (cdf[i + 1] - cdf[i]) * (float)count;
}
__device__ __forceinline__ V3 env_sample(const DScene &S, Rng &rng) { // texture.cpp:79-84
    float u = rng.random();
    float v = rng.random();
    float y = dist_sample(S.env_marg, S.env_h, S.env_gmarg, S.env_gh, v);
    int yi = dist_anchor(S.env_h, y);
    float x = dist_sample(S.env_cond + (size_t)yi * (S.env_w + 1), S.env_w, S.env_gcond + (size_t)yi * (S.env_gw + 1),
                          S.env_gw, u);
    return mat_apply_t(S.bg_rot, uv2dir(V2{x, y}));
}
__device__ __forceinline__ float env_pdf(const DScene &S, V3 dir) { // texture.cpp:86-90
    V2 uv = dir2uv(mat_apply(S.bg_rot, dir));
    float py = dist_pdf(S.env_marg, S.env_h, uv.y);
    int yi = dist_anchor(S.env_h, uv.y);
    float px = dist_pdf(S.env_cond + (size_t)yi * (S.env_w + 1), S.env_w, uv.x);
    return (py * px) * (0.25f * kINV_PI);
}

// ---------------------------------------------------------------- primitive distance tests
// Each returns the reference's candidate t (kINF when none) and, for spheres, `into`.
__device__ __forceinline__ float sphere_t(const DAnalytic &p, V3 o, V3 d, bool &into) { // object.cpp:21-42
    V3 op = sub(mk(p.center), o);
    float b = dot(op, d);
    float det = sq(b) - dot(op, op) + sq(p.radius);
    into = true;
    if (det < 0.f) return kINF;
    det = __builtin_sqrtf(det);
    if (b - det >= kEPS) return b - det;
    if (b + det >= kEPS) {
        into = false;
        return b + det;
    }
    return kINF;
}
// Rectangle / Circle (object.cpp:114-131, 163-180): t and the in-shape test
__device__ __forceinline__ float plane_t(const DAnalytic &p, V3 o, V3 d, float tmax, V2 &uv) {
    V3 n = mk(p.w);
    float t = (p.dist - dot(o, n)) / dot(d, n);
    if (!(kEPS <= t && t < tmax)) return kINF;
    V3 cp = sub(at(o, d, t), mk(p.center));
    if (p.type == A_RECT) {
        float u = dot(cp, mk(p.u)) / p.dim[0] + 0.5f;
        float v = dot(cp, mk(p.v)) / p.dim[1] + 0.5f;
        if (0.f <= u && u <= 1.f && 0.f <= v && v <= 1.f) {
            uv = {u, v};
            return t;
        }
        return kINF;
    }
    if (sqn(cp) <= sq(p.radius)) {
        V3 c = dvs(cp, p.radius);
        uv = {dot(c, mk(p.u)), dot(c, mk(p.v))};
        return t;
    }
    return kINF;
}
// Moller-Trumbore on the reference's stored (a, ab, ac) (object.cpp:252-264)
__device__ __forceinline__ bool tri_test(V3 a, V3 ab, V3 ac, V3 o, V3 d, float &t, float &u, float &v) {
    V3 pv = cross(d, ac);
    float inv_det = 1.f / dot(pv, ab);
    V3 ao = sub(o, a);
    u = dot(pv, ao) * inv_det;
    V3 q = cross(ao, ab);
    v = dot(q, d) * inv_det;
    if (!(0.f <= u && 0.f <= v && u + v <= 1.f)) return false;
    t = dot(q, ac) * inv_det;
    return true;
}

// ---------------------------------------------------------------- light sampling (object.cpp, light.cpp)
__device__ __forceinline__ float cos_theta_max(const DAnalytic &p, V3 org) { // object.cpp:95-104
    float d2 = sqn(sub(mk(p.center), org));
    float r2 = sq(p.radius);
    if (d2 <= r2) return -1.f;
    return __builtin_sqrtf(1.f - r2 / d2);
}
__device__ __forceinline__ float plane_pdf(float dist, float area, V3 n, V3 dir) { // object.cpp:11-16
    float d2 = sq(dist);
    float ca = fabsf(dot(n, dir));
    float pdf = d2 / (ca * area);
    return (pdf < kINF) ? pdf : kINF;
}
__device__ __forceinline__ V3 analytic_sample(const DAnalytic &p, V3 org, Rng &rng) {
    if (p.type == A_SPHERE) { // object.cpp:66-75
        float ctm = cos_theta_max(p, org);
        float ct = rng.random(ctm, 1.f);
        float st = __builtin_sqrtf(1.f - ct * ct);
        float phi = rng.random(2.f * kPI);
        V3 local{st * cosf(phi), st * sinf(phi), ct};
        return frame_apply(from_direction(nrm(sub(mk(p.center), org))), local);
    }
    if (p.type == A_RECT) { // object.cpp:133-138
        V3 uvec = scl((rng.random() - .5f) * p.dim[0], mk(p.u));
        V3 vvec = scl((rng.random() - .5f) * p.dim[1], mk(p.v));
        return nrm(sub(add(add(mk(p.center), uvec), vvec), org));
    }
    V2 uv = rng.rand_in_disk(p.radius); // object.cpp:188-193
    V3 target = add(add(mk(p.center), scr(mk(p.u), uv.x)), scr(mk(p.v), uv.y));
    return nrm(sub(target, org));
}
__device__ __forceinline__ float analytic_pdf(const DAnalytic &p, V3 org, V3 dir) {
    if (p.type == A_SPHERE) { // object.cpp:77-88
        bool into;
        if (!(sphere_t(p, org, dir, into) < kINF)) return 0.f;
        float ctm = cos_theta_max(p, org);
        float sa = 2.f * kPI * (1.f - ctm);
        return 1.f / sa;
    }
    V2 uv;
    float t = plane_t(p, org, dir, kINF, uv);
    if (!(t < kINF)) return 0.f;
    float area = (p.type == A_RECT) ? p.dim[0] * p.dim[1] : kPI * sq(p.radius);
    return plane_pdf(t, area, mk(p.w), dir);
}
__device__ __forceinline__ V3 tri_sample(const DScene &S, int tri, V3 org, Rng &rng) { // object.cpp:226-235
    float4 r0 = S.tris[3 * tri], r1 = S.tris[3 * tri + 1], r2v = S.tris[3 * tri + 2];
    float sr1 = __builtin_sqrtf(rng.random());
    float r2 = rng.random();
    float u = sr1 * (1.f - r2);
    float v = r2 * sr1;
    V3 target = add(add(mk4(r0), scl(u, mk4(r1))), scl(v, mk4(r2v)));
    return nrm(sub(target, org));
}
__device__ __forceinline__ float tri_pdf(const DScene &S, int tri, V3 org, V3 dir) { // object.cpp:237-248
    // scalar loads when tri is wave-uniform (the light loop of lights_pdf)
    float4 r0 = load_invariant(S.tris + 3 * tri), r1 = load_invariant(S.tris + 3 * tri + 1),
           r2v = load_invariant(S.tris + 3 * tri + 2);
    V3 ab = mk4(r1), ac = mk4(r2v);
    float t, u, v;
    if (!tri_test(mk4(r0), ab, ac, org, dir, t, u, v)) return 0.f;
    if (!(kEPS <= t && t < kINF)) return 0.f;
    V3 fn = cross(ac, ab);
    float area = .5f * __builtin_sqrtf(sqn(fn));
    fn = nrm(fn);
    return plane_pdf(t, area, fn, dir);
}
// ObjectLight of a whole mesh: a shape, then a face of it, uniformly, then the
// face's sample; pdf = mean over shapes of the mean over their faces
// (ObjectGroup::sample / pdf, cpu/object.cpp:352-365)
__device__ __forceinline__ V3 mesh_light_sample(const DScene &S, const DLight &l, V3 pos, Rng &rng) {
    const int2 r = S.light_shapes[l.local + rng.rand_int(l.gid)];
    return tri_sample(S, r.x + rng.rand_int(r.y - r.x), pos, rng);
}
__device__ __forceinline__ float mesh_light_pdf(const DScene &S, const DLight &l, V3 pos, V3 dir) {
    float val = 0.f;
    for (int s = 0; s < l.gid; s++) {
        const int2 r = load_invariant(S.light_shapes + l.local + s);
        float sv = 0.f;
        for (int tri = r.x; tri < r.y; tri++) sv += tri_pdf(S, tri, pos, dir);
        val += sv / (float)(r.y - r.x);
    }
    return val / (float)l.gid;
}
__device__ __forceinline__ V3 lights_sample(const DScene &S, V3 pos, Rng &rng) { // LightGroup::sample
    const DLight l = S.lights[rng.rand_int(S.num_lights) ];
    if (l.kind == L_ENV) return env_sample(S, rng);
    if (l.kind == L_ANALYTIC) return analytic_sample(S.analytic[l.local], pos, rng);
    if (l.kind == L_MESH) return mesh_light_sample(S, l, pos, rng);
    return tri_sample(S, l.local, pos, rng);
}
__device__ __forceinline__ float lights_pdf(const DScene &S, V3 pos, V3 dir) { // LightGroup::pdf
    float val = 0.f;
    for (int i = 0; i < S.num_lights; i++) {
        const DLight l = load_invariant(S.lights + i); // i is wave-uniform: scalar loads
        float p;
        if (l.kind == L_ENV) p = env_pdf(S, dir);
        else if (l.kind == L_ANALYTIC) p = analytic_pdf(load_invariant(S.analytic + l.local), pos, dir);
        else if (l.kind == L_MESH) p = mesh_light_pdf(S, l, pos, dir);
        else p = tri_pdf(S, l.local, pos, dir);
        val += p;
    }
    return val / (float)S.num_lights;
}
// DistantLight / PointLight sample (light.cpp:11-27); tmax = light distance
__device__ __forceinline__ V3 delta_sample(const DDelta &l, V3 pos, Rng &rng, float &tmax) {
    if (l.type == 0) {
        tmax = kINF;
        if (l.angular_radius < kEPS) return mk(l.w);
        float du = rng.random(-l.angular_radius, l.angular_radius);
        float dv = rng.random(-l.angular_radius, l.angular_radius);
        Frame f{mk(l.u), mk(l.v), mk(l.w)};
        return nrm(frame_apply(f, V3{du, dv, 1.f}));
    }
    V3 c = mk(l.center);
    if (l.radius >= kEPS) c = add(c, rng.rand_on_sphere(l.radius));
    V3 d = sub(c, pos);
    tmax = __builtin_sqrtf(sqn(d));
    return nrm(d);
}

// ---------------------------------------------------------------- BxDFs (cpu/material.cpp)
struct SurfHit {
    V3 normal, shade_normal;
    V2 uv;
    bool into;
};

__device__ __forceinline__ float micro_D(const DLobe &b, V3 n, V3 h) {
    if (b.type == 3) { // BlinnPhongModel::D_impl (material.cpp:70-73)
        float cb = dot(h, n);
        return (cb <= 0.f) ? 0.f : (b.param + 2.f) * (.5f * kINV_PI) * powf(cb, b.param + 1.f);
    }
    float cb = dot(n, h); // TrowbridgeReitzModel::D_impl (material.cpp:103-111)
    if (cb <= 0.f) return 0.f;
    float c2 = cb * cb;
    float s2 = 1.f - c2;
    return b.alpha2 * kINV_PI * cb / sq(b.alpha2 * c2 + s2);
}
__device__ __forceinline__ float blinn_G1(float shin, V3 w, V3 n) { // material.cpp:41-54
    float ct = dot(w, n);
    ct = fmaxf(fminf(ct, 1.f), 0.f);
    float st = __builtin_sqrtf(1.f - ct * ct);
    float a = __builtin_sqrtf(.5f * shin + 1.f) * ct / st;
    if ((double)a < 1.6) {
        float a2 = a * a;
        return (3.535f * a + 2.181f * a2) / (1.f + 2.276f * a + 2.577f * a2);
    }
    return 1.f;
}
__device__ __forceinline__ float ggx_G1(float alpha2, V3 w, V3 n) { // material.cpp:78-84
    float ct = dot(w, n);
    float tan2 = 1.f / sq(ct) - 1.f;
    return 2.f / (1.f + __builtin_sqrtf(1.f + alpha2 * tan2));
}
__device__ __forceinline__ V3 lobe_f(const DScene &S, const DLobe &b, V3 rd, const SurfHit &h, V3 out) {
    if (b.type == 0) { // Lambertian::f (material.cpp:6-13)
        float c = dot(h.shade_normal, out);
        if (c > 0.f) return scl(c * kINV_PI, color_at(S, b.color, h.uv));
        return V3{0.f, 0.f, 0.f};
    }
    if (b.type == 1 || b.type == 2) return color_at(S, b.color, h.uv); // Metal / Dielectric ::f
    // Glossy::f (material.cpp:123-134)
    V3 l = neg(rd);
    V3 n = h.shade_normal;
    if (!same_side(out, n)) return V3{0.f, 0.f, 0.f};
    V3 hv = nrm(add(out, l));
    V3 col = color_at(S, b.color, h.uv);
    float D = micro_D(b, n, hv);
    float G = (b.type == 3) ? blinn_G1(b.param, l, n) * blinn_G1(b.param, out, n)
                            : ggx_G1(b.alpha2, l, n) * ggx_G1(b.alpha2, out, n);
    float s = D * G / (4.f * dot(n, l) * dot(n, out));
    float vh = 1.f - dot(out, hv);
    float p5 = (vh * vh) * (vh * vh) * vh;
    V3 F{col.x + (1.f - col.x) * p5, col.y + (1.f - col.y) * p5, col.z + (1.f - col.z) * p5};
    V3 d = scr(F, s);
    return finite3(d) ? d : V3{0.f, 0.f, 0.f};
}
__device__ __forceinline__ V3 lambert_sample(const SurfHit &h, Rng &rng) { // material.cpp:23-33
    float phi = rng.random(2.f * kPI);
    float s2 = rng.random();
    float st = __builtin_sqrtf(s2);
    float ct = __builtin_sqrtf(1.f - s2);
    return frame_apply(from_direction(h.normal), V3{cosf(phi) * st, sinf(phi) * st, ct});
}
__device__ __forceinline__ V3 lobe_sample(const DLobe &b, V3 rd, const SurfHit &h, Rng &rng) {
    if (b.type == 0) return lambert_sample(h, rng);
    if (b.type == 1) return reflect(rd, h.normal); // Metal::sample
    if (b.type == 2) {                             // Dielectric::sample (material.cpp:179-208)
        V3 refl = reflect(rd, h.normal);
        const float nc = 1.f, ng = b.param;
        float eta = h.into ? (nc / ng) : (ng / nc);
        float cosi = -dot(rd, h.normal);
        float cos2t = 1.f - eta * eta * (1.f - cosi * cosi);
        if (cos2t < 0.f) return refl;
        float cost = __builtin_sqrtf(cos2t);
        V3 refr = nrm(add(scl(eta, rd), scl(eta * cosi - cost, h.normal)));
        float R0 = sq(ng - nc) / sq(ng + nc);
        float c = 1.f - (h.into ? cosi : cost);
        float Re = R0 + (1.f - R0) * c * c * c * c * c;
        return (rng.random() < Re) ? refl : refr;
    }
    float cb, sb, phi; // Glossy::sample with sample_h (material.cpp:58-66, 88-99, 136-144)
    if (b.type == 3) {
        cb = powf(rng.random(), 1.f / (b.param + 2.f));
        sb = __builtin_sqrtf(1.f - cb * cb);
        phi = rng.random(2.f * kPI);
    } else {
        float u = rng.random();
        float c2 = (1.f - u) / (1.f - (1.f - b.alpha2) * u);
        cb = __builtin_sqrtf(c2);
        sb = __builtin_sqrtf(1.f - c2);
        phi = rng.random(2.f * kPI);
    }
    V3 hv = frame_apply(from_direction(h.normal), V3{cosf(phi) * sb, sinf(phi) * sb, cb});
    V3 out = reflect(rd, hv);
    if (!same_side(out, h.normal)) return lambert_sample(h, rng);
    return out;
}
__device__ __forceinline__ float lobe_pdf(const DLobe &b, V3 rd, const SurfHit &h, V3 out) {
    if (b.type == 0) { // Lambertian::pdf
        float c = dot(out, h.normal);
        return (c > 0.f) ? c * kINV_PI : 0.f;
    }
    V3 l = neg(rd); // Glossy::pdf (material.cpp:146-155)
    V3 n = h.normal;
    if (!same_side(out, n)) return 0.f;
    V3 hv = nrm(add(out, l));
    return micro_D(b, n, hv) / (4.f * dot(n, l) * dot(n, out));
}
// sel = -1: the combined non-specular Add/Mix surface
__device__ __forceinline__ V3 bxdf_f(const DScene &S, const DMaterial &m, int sel, V3 rd, const SurfHit &h, V3 out) {
    if (sel >= 0) return lobe_f(S, m.lobe[sel], rd, h, out);
    V3 f0 = lobe_f(S, m.lobe[0], rd, h, out), f1 = lobe_f(S, m.lobe[1], rd, h, out);
    if (m.combine == 1) return add(f0, f1);
    return add(scl(m.weight, f0), scl(1.f - m.weight, f1));
}
__device__ __forceinline__ V3 bxdf_sample(const DMaterial &m, int sel, V3 rd, const SurfHit &h, Rng &rng) {
    if (sel >= 0) return lobe_sample(m.lobe[sel], rd, h, rng);
    float thr = (m.combine == 1) ? .5f : m.weight;
    return (rng.random() < thr) ? lobe_sample(m.lobe[0], rd, h, rng) : lobe_sample(m.lobe[1], rd, h, rng);
}
__device__ __forceinline__ float bxdf_pdf(const DMaterial &m, int sel, V3 rd, const SurfHit &h, V3 out) {
    if (sel >= 0) return lobe_pdf(m.lobe[sel], rd, h, out);
    float p0 = lobe_pdf(m.lobe[0], rd, h, out), p1 = lobe_pdf(m.lobe[1], rd, h, out);
    if (m.combine == 1) return .5f * (p0 + p1);
    return m.weight * p0 + (1.f - m.weight) * p1;
}
__device__ __forceinline__ int sample_bxdf(const DMaterial &m, Rng &rng, float &pdf_bxdf, bool &spec) {
    if (m.combine == 0) {
        pdf_bxdf = 1.f;
        spec = m.lobe[0].specular != 0;
        return 0;
    }
    if (!m.specular) {
        pdf_bxdf = 1.f;
        spec = false;
        return -1;
    }
    int sel;
    if (m.combine == 1) {
        sel = (rng.random() < .5f) ? 0 : 1;
        pdf_bxdf = 0.5f;
    } else {
        sel = (rng.random() < m.weight) ? 0 : 1;
        pdf_bxdf = sel == 0 ? m.weight : 1.f - m.weight;
    }
    spec = m.lobe[sel].specular != 0;
    return sel;
}

} // namespace tpt
