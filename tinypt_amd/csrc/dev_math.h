// dev_math.h — gfx950 device math for the path tracer.
//
// Built with -ffp-contract=off and correctly rounded f32 divide/sqrt, so each
// + - * / sqrt below is exactly one IEEE op; the summation orders restate the
// Eigen expressions of the reference CPU path (li-plus/tinypt src/tinypt/cpu/):
//   dot(a, b) = a0 b0 + (a1 b1 + a2 b2)   (Eigen fixed-size redux)
//   normalized(v) = v / sqrt(v.v)
//   Rotation * v = (u v0 + v v1) + w v2   (cpu/mapping.h:34)
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "dev_types.h"

namespace tpt {

static constexpr float kPI = 3.14159265358979323846f;      // M_PIf32
static constexpr float kINV_PI = 0.318309886183790671538f; // M_1_PIf32

struct V3 {
    float x, y, z;
};
struct V2 {
    float x, y;
};

__device__ __forceinline__ V3 mk(float x, float y, float z) { return V3{x, y, z}; }
__device__ __forceinline__ V3 mk(const float *p) { return V3{p[0], p[1], p[2]}; }
__device__ __forceinline__ V3 mk4(float4 a) { return V3{a.x, a.y, a.z}; }
__device__ __forceinline__ V3 add(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ V3 neg(V3 a) { return {-a.x, -a.y, -a.z}; }
__device__ __forceinline__ V3 scl(float s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
__device__ __forceinline__ V3 scr(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ V3 dvs(V3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
__device__ __forceinline__ V3 cmul(V3 a, V3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
__device__ __forceinline__ float dot(V3 a, V3 b) { return a.x * b.x + (a.y * b.y + a.z * b.z); }
__device__ __forceinline__ float sqn(V3 a) { return dot(a, a); }
__device__ __forceinline__ V3 cross(V3 a, V3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ V3 nrm(V3 a) {
    float n = sqn(a);
    if (n > 0.f) {
        float s = __builtin_sqrtf(n);
        return dvs(a, s);
    }
    return a;
}
__device__ __forceinline__ float sq(float x) { return x * x; }
__device__ __forceinline__ bool same_side(V3 a, V3 b) { return dot(a, b) > 0.f; }
__device__ __forceinline__ bool is_zero5(V3 a) {
    return fabsf(a.x) <= 1e-5f && fabsf(a.y) <= 1e-5f && fabsf(a.z) <= 1e-5f;
}
__device__ __forceinline__ bool finite3(V3 a) { return isfinite(a.x) && isfinite(a.y) && isfinite(a.z); }
__device__ __forceinline__ float clip1(float v) {
    float m = (v < 0.f) ? 0.f : v;
    return (1.f < m) ? 1.f : m;
}
__device__ __forceinline__ V3 clip01(V3 a) { return {clip1(a.x), clip1(a.y), clip1(a.z)}; }
__device__ __forceinline__ V3 at(V3 o, V3 d, float t) { return add(o, scr(d, t)); }

struct Frame {
    V3 u, v, w;
};
__device__ __forceinline__ V3 frame_apply(const Frame &f, V3 a) { // Rotation::operator*
    return add(add(scr(f.u, a.x), scr(f.v, a.y)), scr(f.w, a.z));
}
__device__ __forceinline__ Frame from_direction(V3 w) { // Rotation::from_direction, mapping.cpp:20-30
    V3 u;
    if (fabsf(w.y) < .9f) {
        u = nrm(V3{w.z, 0.f, -w.x});
    } else {
        u = nrm(V3{0.f, -w.z, w.y});
    }
    return Frame{u, cross(w, u), w};
}
// Mat3f (columns u,v,w) * a, row-wise dot order
__device__ __forceinline__ V3 mat_apply(const float *m, V3 a) {
    return {m[0] * a.x + (m[3] * a.y + m[6] * a.z), m[1] * a.x + (m[4] * a.y + m[7] * a.z),
            m[2] * a.x + (m[5] * a.y + m[8] * a.z)};
}
__device__ __forceinline__ V3 mat_apply_t(const float *m, V3 a) {
    return {dot(mk(m), a), dot(mk(m + 3), a), dot(mk(m + 6), a)};
}
__device__ __forceinline__ V3 reflect(V3 d, V3 n) { return sub(d, scl(2.f * dot(n, d), n)); }

// ---- erand48 / rand_r stream (cpu/rand.h:10-50), state = 48-bit LCG value
struct Rng {
    uint64_t x;
    __device__ __forceinline__ float random() {
        x = (x * 0x5DEECE66Dull + 0xBull) & 0xFFFFFFFFFFFFull;
        return (float)((double)x * 0x1p-48);
    }
    __device__ __forceinline__ float random(float hi) { return hi * random(); }
    __device__ __forceinline__ float random(float lo, float hi) { return (hi - lo) * random() + lo; }
    __device__ __forceinline__ int rand_int() {
        uint32_t next = (uint32_t)(x & 0xFFFFFFFFull);
        uint32_t r;
        next = next * 1103515245u + 12345u;
        r = (next / 65536u) % 2048u;
        next = next * 1103515245u + 12345u;
        r <<= 10;
        r ^= (next / 65536u) % 1024u;
        next = next * 1103515245u + 12345u;
        r <<= 10;
        r ^= (next / 65536u) % 1024u;
        x = (x & 0xFFFF00000000ull) | (uint64_t)next;
        return (int)r;
    }
    __device__ __forceinline__ int rand_int(int n) { return rand_int() % n; } // rand.h:24
    __device__ __forceinline__ V2 rand_in_disk(float radius) {
        float th = random(2.f * kPI);
        float c = cosf(th), s = sinf(th);
        float r = __builtin_sqrtf(random());
        return {(c * r) * radius, (s * r) * radius};
    }
    __device__ __forceinline__ V3 rand_on_sphere(float radius) {
        float phi = random(2.f * kPI);
        float ct = random(-1.f, 1.f);
        float st = __builtin_sqrtf(1.f - ct * ct);
        return scr(V3{st * cosf(phi), st * sinf(phi), ct}, radius);
    }
};
__host__ __device__ __forceinline__ uint64_t rng_state_from_seed(uint64_t seed) { // RandEngine(uint64_t)
    uint64_t s0 = seed & 0xffff, s1 = (seed >> 16) & 0xffff;
    uint64_t s2 = (((seed >> 32) & 0xffff) + ((seed >> 48) & 0xffff)) & 0xffff;
    return (s2 << 32) | (s1 << 16) | s0;
}
__host__ __device__ __forceinline__ uint64_t splitmix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
// one stream per (pixel, sample)
__host__ __device__ __forceinline__ uint64_t stream_seed(uint64_t base_mixed, uint64_t pixel, uint64_t sample) {
    return splitmix64(base_mixed ^ (pixel << 24) ^ sample);
}
// order-independent alpha cut-out draw per (path, depth, primitive)
__device__ __forceinline__ float alpha_draw(uint64_t key, int depth, int32_t prim) {
    uint64_t h = splitmix64(key ^ ((uint64_t)(uint32_t)depth << 40) ^ ((uint64_t)(uint32_t)prim * 0x9E3779B97F4A7C15ull));
    return (float)(h >> 40) * 0x1p-24f;
}

} // namespace tpt
