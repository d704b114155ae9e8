// oracle_rng.h — TEST INFRASTRUCTURE ONLY (the parity oracle).
//
// Restatement of cpu::RandEngine (src/tinypt/cpu/rand.h:10-50): a glibc erand48
// float stream and a glibc rand_r int stream aliased on the same 48-bit state.
//   erand48: POSIX 48-bit LCG, a = 0x5DEECE66D, c = 0xB; glibc returns the NEW
//            state X as the double X / 2^48 (exact), cast to float by the caller
//            (so random() can return 1.0f).
//   rand_r : glibc's three-step LCG (1103515245, 12345) on the first 4 bytes of
//            the state (rand.h:23 casts uint16_t[3] to uint32_t*), little-endian.
// Pinned by known-answer vectors generated from the system glibc (tests/golden/).
#pragma once

#include <cmath>
#include <cstdint>

#include "oracle_math.h"

namespace oracle {

struct RandEngine {
    uint16_t s[3];

    // RandEngine(uint64_t seed_) (rand.h:13-17)
    explicit RandEngine(uint64_t seed) {
        s[0] = (uint16_t)(seed & 0xffff);
        s[1] = (uint16_t)((seed >> 16) & 0xffff);
        s[2] = (uint16_t)(((seed >> 32) & 0xffff) + ((seed >> 48) & 0xffff));
    }

    double erand48() {
        uint64_t x = ((uint64_t)s[2] << 32) | ((uint64_t)s[1] << 16) | (uint64_t)s[0];
        x = (x * 0x5DEECE66Dull + 0xBull) & 0xFFFFFFFFFFFFull;
        s[0] = (uint16_t)(x & 0xffff);
        s[1] = (uint16_t)((x >> 16) & 0xffff);
        s[2] = (uint16_t)((x >> 32) & 0xffff);
        return (double)x * 0x1p-48;
    }
    float random() { return (float)erand48(); }                                 // rand.h:19
    float random(float hi) { return hi * random(); }                           // rand.h:20
    float random(float lo, float hi) { return (hi - lo) * random() + lo; }      // rand.h:21

    int rand_int() { // rand.h:23
        uint32_t next = (uint32_t)s[0] | ((uint32_t)s[1] << 16);
        uint32_t result;
        next = next * 1103515245u + 12345u;
        result = (next / 65536u) % 2048u;
        next = next * 1103515245u + 12345u;
        result <<= 10;
        result ^= (next / 65536u) % 1024u;
        next = next * 1103515245u + 12345u;
        result <<= 10;
        result ^= (next / 65536u) % 1024u;
        s[0] = (uint16_t)(next & 0xffff);
        s[1] = (uint16_t)(next >> 16);
        return (int)result;
    }
    int rand_int(int n) { return rand_int() % n; } // rand.h:24

    // rand_on_disk / rand_in_disk (rand.h:27-34).  The reference writes
    // `rand_on_disk() * std::sqrt(random())`, whose operand order C++ leaves
    // unspecified; the oracle fixes it left to right (angle first).
    V2 rand_on_disk() {
        float theta = random(2.f * PI_F);
        return {std::cos(theta), std::sin(theta)};
    }
    V2 rand_in_disk(float radius) {
        V2 d = rand_on_disk();
        float r = std::sqrt(random());
        return {(d.x * r) * radius, (d.y * r) * radius};
    }
    // rand_on_sphere (rand.h:36-43)
    V3 rand_on_sphere(float radius) {
        float phi = random(2.f * PI_F);
        float cos_theta = random(-1, 1);
        float sin_theta = std::sqrt(1 - cos_theta * cos_theta);
        V3 p{sin_theta * std::cos(phi), sin_theta * std::sin(phi), cos_theta};
        return p * radius;
    }
};

// Per-(pixel, sample) stream seeds.  The reference seeds one stream per image
// row from the wall clock (cpu/path_tracer.cpp:65-66,76), which is neither
// deterministic nor parallel; this build seeds one stream per (pixel, sample)
// with splitmix64 so that any scheduling reproduces the same paths.
static inline uint64_t splitmix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
static inline uint64_t stream_seed(uint64_t base, uint64_t pixel, uint64_t sample) {
    return splitmix64(splitmix64(base) ^ (pixel << 24) ^ sample);
}
// Order-independent alpha cut-out draw (replaces the in-traversal rng draw of
// cpu/object.cpp:53,285, whose count depends on traversal order).
static inline float alpha_draw(uint64_t path_key, int depth, int32_t prim) {
    uint64_t h = splitmix64(path_key ^ ((uint64_t)(uint32_t)depth << 40) ^ ((uint64_t)(uint32_t)prim * 0x9E3779B97F4A7C15ull));
    return (float)(h >> 40) * 0x1p-24f;
}

} // namespace oracle
