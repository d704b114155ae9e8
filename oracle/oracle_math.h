// oracle_math.h — TEST INFRASTRUCTURE ONLY (the parity oracle; never shipped or linked by the product).
//
// Small fixed-size vector math restating the evaluation order of the Eigen
// expressions used by the reference CPU path (src/tinypt/cpu/defs.h:11-21).
// Compiled with -ffp-contract=off so every + - * / sqrt is one IEEE op.
//   dot / squaredNorm : Eigen fixed-size redux = x0 + (x1 + x2)      (parity unpinned at ulp level:
//                       Eigen is absent here; order restated from Eigen 3.3's redux_novec_unroller)
//   normalized()      : v / sqrt(v.v) when v.v > 0, else v
//   Mat3f * Vec3f     : row i = m(i,0) v0 + (m(i,1) v1 + m(i,2) v2)  (coefficient-based product)
//   Rotation * Vec3f  : (u x + v y) + w z                              (cpu/mapping.h:34)
#pragma once

#include <cmath>
#include <cstdint>

namespace oracle {

static constexpr float INF = 1e20f;            // cpu/defs.h:23
static constexpr float EPS = 1e-3f;            // cpu/defs.h:24
static constexpr float PI_F = 3.14159265358979323846f;   // M_PIf32
static constexpr float INV_PI_F = 0.318309886183790671538f; // M_1_PIf32

struct V3 {
    float x, y, z;
    float operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
    float &operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
};
struct V2 {
    float x, y;
};

static inline V3 v3(float x, float y, float z) { return V3{x, y, z}; }
static inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
static inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
static inline V3 operator-(V3 a) { return {-a.x, -a.y, -a.z}; }
static inline V3 operator*(float s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
static inline V3 operator*(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
static inline V3 operator/(V3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
static inline V3 cmul(V3 a, V3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
static inline float dot(V3 a, V3 b) { return a.x * b.x + (a.y * b.y + a.z * b.z); }
static inline float sqnorm(V3 a) { return dot(a, a); }
static inline float norm(V3 a) { return std::sqrt(sqnorm(a)); }
static inline V3 cross(V3 a, V3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
static inline V3 normalized(V3 a) {
    float n = sqnorm(a);
    if (n > 0) {
        float s = std::sqrt(n);
        return a / s;
    }
    return a;
}
static inline float square(float x) { return x * x; }
static inline bool same_side(V3 a, V3 b) { return dot(a, b) > 0; } // cpu/defs.h:33
static inline V3 cwise_min(V3 a, V3 b) { return {a.x < b.x ? a.x : b.x, a.y < b.y ? a.y : b.y, a.z < b.z ? a.z : b.z}; }
static inline V3 cwise_max(V3 a, V3 b) { return {a.x > b.x ? a.x : b.x, a.y > b.y ? a.y : b.y, a.z > b.z ? a.z : b.z}; }
// Vec3f::isZero() with Eigen's float dummy_precision 1e-5 (cpu/path_tracer.cpp:161)
static inline bool is_zero(V3 a) {
    return std::fabs(a.x) <= 1e-5f && std::fabs(a.y) <= 1e-5f && std::fabs(a.z) <= 1e-5f;
}
static inline bool all_finite(V3 a) { return std::isfinite(a.x) && std::isfinite(a.y) && std::isfinite(a.z); }
static inline V3 clip01(V3 a) { // cpu/path_tracer.cpp:60: cwiseMax(0) then cwiseMin(1), std::max/min semantics
    auto c = [](float v) {
        float m = (v < 0.f) ? 0.f : v;
        return (1.f < m) ? 1.f : m;
    };
    return {c(a.x), c(a.y), c(a.z)};
}

// 3x3 stored column-major as u, v, w columns: m[0..2]=u, m[3..5]=v, m[6..8]=w.
struct Rot {
    V3 u, v, w;
};
static inline Rot rot_from9(const float *m) { return Rot{{m[0], m[1], m[2]}, {m[3], m[4], m[5]}, {m[6], m[7], m[8]}}; }
// Rotation::operator* (cpu/mapping.h:34)
static inline V3 rot_apply(const Rot &r, V3 a) { return (r.u * a.x + r.v * a.y) + r.w * a.z; }
// Mat3f * Vec3f: row i of the matrix with columns u,v,w is (u[i], v[i], w[i]).
static inline V3 mat_apply(const Rot &r, V3 a) {
    return {r.u.x * a.x + (r.v.x * a.y + r.w.x * a.z), r.u.y * a.x + (r.v.y * a.y + r.w.y * a.z),
            r.u.z * a.x + (r.v.z * a.y + r.w.z * a.z)};
}
// Mat3f^T * Vec3f (inverse of an orthonormal frame).
static inline V3 mat_apply_t(const Rot &r, V3 a) { return {dot(r.u, a), dot(r.v, a), dot(r.w, a)}; }

// Rotation::from_direction (cpu/mapping.cpp:20-30)
static inline Rot from_direction(V3 w) {
    V3 u;
    if (std::fabs(w.y) < .9f) {
        u = normalized(V3{w.z, 0.f, -w.x});
    } else {
        u = normalized(V3{0.f, -w.z, w.y});
    }
    V3 v = cross(w, u);
    return Rot{u, v, w};
}

// Ray::reflect (cpu/ray.h:19-24) and Ray::refract (cpu/ray.h:26-29)
static inline V3 reflect(V3 d, V3 n) { return d - (2.f * dot(n, d)) * n; }
static inline V3 refract(V3 d, V3 n, float eta, float cos_i, float cos_t) {
    return normalized(eta * d + (eta * cos_i - cos_t) * n);
}

} // namespace oracle
