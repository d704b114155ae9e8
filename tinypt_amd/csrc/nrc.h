// nrc.h — learned radiance cache: weight storage, the 64-dimensional input
// encoding used by the shade kernel, and the MLP kernel launchers (nrc.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "dev_math.h"

namespace tpt {

// FP32 -> BF16 bits, round to nearest even (NaN kept quiet).
__host__ __device__ inline uint16_t f32_to_bf16_bits(float f) {
    uint32_t u = __builtin_bit_cast(uint32_t, f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

// Two FP32 values rounded to BF16 (to nearest even) by one v_cvt_pk_bf16_f32,
// `a` in the low half; the same bits as f32_to_bf16_bits for finite values.
typedef __attribute__((ext_vector_type(2))) __bf16 bf16x2_t;
typedef __attribute__((ext_vector_type(2))) float f32x2_t;
__device__ inline uint32_t bf16_pair(float a, float b) {
    return __builtin_bit_cast(uint32_t, __builtin_convertvector(f32x2_t{a, b}, bf16x2_t));
}

// Device buffers of the MLP: 5 layers x 64 x 64 (the last layer uses 3 outputs).
struct NrcWeights {
    float *w;      // FP32 master weights [l][in][out]
    float *m, *v;  // Adam moments
    float *grad;   // gradient of the latest step (written by the Adam kernel)
    float *partial; // per-workgroup gradients of a training step (kNrcMaxTrainGroups x kNrcParams)
    uint16_t *wt;  // BF16 [l][out][in] (forward)
    uint16_t *wb;  // BF16 [l][in][out] (backward)
    float *loss;   // running sum of the per-step mean loss
};
static constexpr int kNrcParams = 5 * 64 * 64;
// Training workgroups per step at most (each keeps its own gradient; larger
// steps give a workgroup several 64-record tiles).
static constexpr int kNrcMaxTrainGroups = 256;
static constexpr int kNrcInputs = 64;

// Input encoding (64 values): position normalised to the scene bounds with
// sin / cos at kNrcPosOctaves octaves per axis, the outgoing direction and the
// shading normal at kNrcDirOctaves / kNrcNrmOctaves octaves per axis, roughness
// (1) and albedo (3).  Textured room 1920x1080, 256 frames: octaves (6, 2, 2)
// give relMSE 0.0181, (8, 1, 1) 0.0193, (7, 2, 1) 0.0182, (7, 1, 2) 0.0185.
#ifndef TPT_NRC_POS_OCTAVES
#define TPT_NRC_POS_OCTAVES 6
#endif
#ifndef TPT_NRC_DIR_OCTAVES
#define TPT_NRC_DIR_OCTAVES 2
#endif
#ifndef TPT_NRC_NRM_OCTAVES
#define TPT_NRC_NRM_OCTAVES 2
#endif
static constexpr int kNrcPosOctaves = TPT_NRC_POS_OCTAVES, kNrcDirOctaves = TPT_NRC_DIR_OCTAVES,
                     kNrcNrmOctaves = TPT_NRC_NRM_OCTAVES;
static_assert(6 * (kNrcPosOctaves + kNrcDirOctaves + kNrcNrmOctaves) == 60, "the encoding has 64 inputs");
__device__ inline void nrc_encode(uint16_t *dst, V3 pos01, V3 wo, V3 n, float rough, V3 albedo) {
    float f[64];
    int k = 0;
    auto octaves = [&](float v, int count) {
        float x = v * 3.14159265f;
        for (int o = 0; o < count; o++) {
            __sincosf(x, &f[k], &f[k + 1]);
            k += 2;
            x *= 2.f;
        }
    };
    octaves(pos01.x, kNrcPosOctaves);
    octaves(pos01.y, kNrcPosOctaves);
    octaves(pos01.z, kNrcPosOctaves);
    octaves(wo.x, kNrcDirOctaves);
    octaves(wo.y, kNrcDirOctaves);
    octaves(wo.z, kNrcDirOctaves);
    octaves(n.x, kNrcNrmOctaves);
    octaves(n.y, kNrcNrmOctaves);
    octaves(n.z, kNrcNrmOctaves);
    f[60] = rough;
    f[61] = albedo.x;
    f[62] = albedo.y;
    f[63] = albedo.z;
    uint4 *d4 = reinterpret_cast<uint4 *>(dst);
#pragma unroll
    for (int i = 0; i < 8; i++)
        d4[i] = make_uint4(bf16_pair(f[8 * i], f[8 * i + 1]), bf16_pair(f[8 * i + 2], f[8 * i + 3]),
                           bf16_pair(f[8 * i + 4], f[8 * i + 5]), bf16_pair(f[8 * i + 6], f[8 * i + 7]));
}

// feats: records of 64 BF16; `idx` (may be null) maps record i to row idx[i] of
// feats; the record count is read from *n_dev (at most n_max).
void nrc_launch_infer(const uint16_t *feats, const int32_t *idx, const uint32_t *n_dev, int n_max,
                      const uint16_t *wt, float4 *out, int num_cu, hipStream_t st);
// Trains on records [base, base + n_max) of the first min(*n_dev, cap) records
// (the only ones stored: the producer may count more).
void nrc_launch_train(const uint16_t *feats, const int32_t *idx, const float4 *target, const uint32_t *n_dev,
                      int cap, int base, int n_max, const NrcWeights &W, hipStream_t st);
// Adam step on the gradient of the latest nrc_launch_train of n_max records.
void nrc_launch_adam(const NrcWeights &W, int n_max, float lr, int step, hipStream_t st);

} // namespace tpt
