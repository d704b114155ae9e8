// nrc.hip — the learned radiance cache: a 5-layer, 64-wide ReLU MLP (no biases)
// evaluated and trained with gfx950 BF16 MFMA (v_mfma_f32_16x16x32_bf16, FP32
// accumulate).  No counterpart in the reference renderer; it terminates paths of
// the wavefront tracer into the cache after a fixed number of bounces.
//
// Layers: 64 -> 64 -> 64 -> 64 -> 64 -> 3 (the last layer is stored padded to 64
// outputs; its padding columns stay zero).  Parameters are kept as FP32 master
// weights W[l][in][out] (Adam state beside them) plus two BF16 copies:
//   wt[l][out][in]  forward operand  (B[k = in][n = out], 8 consecutive k per lane)
//   wb[l][in][out]  backward operand (B[k = out][n = in], 8 consecutive k per lane)
// so every MFMA B operand is one 16-byte load.
//
// MFMA 16x16x32 lane layout (wave64, lane l, j = 0..7 / 0..3):
//   A[m = l % 16][k = (l / 16) * 8 + j]     B[k = (l / 16) * 8 + j][n = l % 16]
//   D[m = (l / 16) * 4 + j][n = l % 16]
// Inference keeps a tile's activations in registers (k_nrc_infer); training
// stages a 64-sample tile's activations in LDS as [sample][feature] BF16 rows
// (stride kRow) and splits its 16 x 16 output blocks over four wavefronts.
#include <hip/hip_runtime.h>

#include <stdint.h>

#include <algorithm>

#include "nrc.h"

namespace tpt {

typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(4))) float f32x4;

static constexpr int kRow = 72; // LDS row stride in BF16 elements (64 + 8: rows start on different banks)

__device__ __forceinline__ uint16_t f2bf(float f) { return f32_to_bf16_bits(f); }
__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

// 8 consecutive BF16 values as an MFMA operand
__device__ __forceinline__ bf16x8 ld8(const uint16_t *p) { return *reinterpret_cast<const bf16x8 *>(p); }
__device__ __forceinline__ bf16x8 ld8_lds(const uint16_t *p) { return *reinterpret_cast<const bf16x8 *>(p); }
// 8 BF16 values with stride `stride` elements (transposed operand gathers)
__device__ __forceinline__ bf16x8 gather8(const uint16_t *p, int stride) {
    uint16_t v[8];
#pragma unroll
    for (int j = 0; j < 8; j++) v[j] = p[j * stride];
    bf16x8 r;
    __builtin_memcpy(&r, v, 16);
    return r;
}

__device__ __forceinline__ f32x4 mfma(bf16x8 a, bf16x8 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// ---------------------------------------------------------------- inference
// Record count (clamped to n_max) and the feature row of record i.
__device__ __forceinline__ int rec_count(const uint32_t *n_dev, int n_max) { return min((int)*n_dev, n_max); }
__device__ __forceinline__ const uint16_t *rec_row(const uint16_t *feats, const int32_t *idx, int i) {
    return feats + (size_t)(idx ? idx[i] : i) * 64;
}
// out[i]: predicted radiance (rgb, 0) of record i.
//
// The activations stay in registers: each layer computes y^T = W^T x^T with A =
// wt[l] (rows = outputs, 8 K values per lane) and B = x^T (8 K values of one
// sample per lane).  The MFMA result D holds, per lane, 4 consecutive outputs
// (g * 4 + j within a 16-output tile, g = lane / 16) of sample lane % 16, which
// after ReLU and BF16 rounding is the next layer's B operand if the K order of
// step s is permuted: lane group g holds features s * 32 + g * 4 + {0..3} (from
// output tile 2s) and s * 32 + 16 + g * 4 + {0..3} (tile 2s + 1).  The A operand
// reads the same features of a wt row (two 8-byte runs).  The first layer's B
// comes straight from the feature rows in memory (identity K order).  No LDS and
// no barriers: a wavefront owns its 64-record tile from load to store.
__device__ __forceinline__ bf16x8 ld4x2(const uint16_t *p0, const uint16_t *p1) {
    uint2 a = *reinterpret_cast<const uint2 *>(p0), b = *reinterpret_cast<const uint2 *>(p1);
    uint4 v = make_uint4(a.x, a.y, b.x, b.y);
    bf16x8 r;
    __builtin_memcpy(&r, &v, 16);
    return r;
}

// Relu and BF16 rounding (to nearest even, v_cvt_pk_bf16_f32) of 2 x 4 FP32
// values, packed as one B operand.  The ReLU runs on the packed BF16 pairs as a
// signed 16-bit max with 0 (v_pk_max_i16): a BF16 value with the sign bit set
// reads as a negative integer, so every negative value and -0 become +0, the
// same as rounding max(v, 0).
typedef __attribute__((ext_vector_type(2))) short i16x2;
__device__ __forceinline__ uint32_t relu_bf16x2(float a, float b) {
    const i16x2 m = __builtin_elementwise_max(__builtin_bit_cast(i16x2, bf16_pair(a, b)), i16x2{0, 0});
    return __builtin_bit_cast(uint32_t, m);
}
__device__ __forceinline__ bf16x8 relu_pack(f32x4 lo, f32x4 hi) {
    const uint4 v = make_uint4(relu_bf16x2(lo[0], lo[1]), relu_bf16x2(lo[2], lo[3]), relu_bf16x2(hi[0], hi[1]),
                               relu_bf16x2(hi[2], hi[3]));
    return __builtin_bit_cast(bf16x8, v);
}

// Persistent inference: every wavefront loads the network's A operands once (the
// four hidden layers' 4 x 2 tiles and the output layer's 2, 136 registers) and
// then walks 16 * NT-record tiles of the records with a grid stride, so the
// weights are read once per wavefront instead of once per tile.
template <int NT>
__global__ __launch_bounds__(256) void k_nrc_infer(const uint16_t *__restrict__ feats, const int32_t *__restrict__ idx,
                                                   const uint32_t *__restrict__ n_dev, int n_max,
                                                   const uint16_t *__restrict__ wt, float4 *__restrict__ out) {
    const int n = rec_count(n_dev, n_max);
    const int lane = threadIdx.x & 63, g = lane >> 4, r = lane & 15;
    const int wave = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const int waves = gridDim.x * (blockDim.x >> 6);
    if (wave * (NT * 16) >= n) return; // wavefront-uniform (no barriers below)
    // A operands: wt rows (outputs) mt * 16 + r; layer 0 reads features s * 32 + g * 8
    // + {0..7}, the later layers the permuted order of their register-held inputs
    bf16x8 A[4][4][2], A4[2];
#pragma unroll
    for (int ly = 0; ly < 4; ly++)
#pragma unroll
        for (int mt = 0; mt < 4; mt++)
#pragma unroll
            for (int s = 0; s < 2; s++) {
                const uint16_t *row = wt + ly * 4096 + (mt * 16 + r) * 64 + s * 32;
                A[ly][mt][s] = ly == 0 ? ld8(row + g * 8) : ld4x2(row + g * 4, row + 16 + g * 4);
            }
#pragma unroll
    for (int s = 0; s < 2; s++) {
        const uint16_t *row = wt + 4 * 4096 + r * 64 + s * 32;
        A4[s] = ld4x2(row + g * 4, row + 16 + g * 4);
    }
    // software pipeline over the wavefront's tiles: the record rows of the next
    // tile and the row indices of the one after are loaded while a tile computes
    const int stride = waves * (NT * 16);
    auto row_of = [&](int tile0, int nt) { // row index of sample nt * 16 + r of a tile (-1 beyond n)
        const int i = tile0 + nt * 16 + r;
        return i < n ? (idx ? idx[i] : i) : -1;
    };
    auto load_x = [&](const int rows[NT], bf16x8 x[NT][2]) { // layer 0 operand of a tile
#pragma unroll
        for (int nt = 0; nt < NT; nt++) {
            const uint16_t *row = feats + (size_t)(rows[nt] < 0 ? 0 : rows[nt]) * 64;
#pragma unroll
            for (int s = 0; s < 2; s++) {
                x[nt][s] = ld8(row + s * 32 + g * 8);
                if (rows[nt] < 0) x[nt][s] = bf16x8{};
            }
        }
    };
    int rows_next[NT];
    bf16x8 x_next[NT][2];
    {
        int rows0[NT];
#pragma unroll
        for (int nt = 0; nt < NT; nt++) rows0[nt] = row_of(wave * (NT * 16), nt);
        load_x(rows0, x_next);
#pragma unroll
        for (int nt = 0; nt < NT; nt++) rows_next[nt] = row_of(wave * (NT * 16) + stride, nt);
    }
    for (int tile0 = wave * (NT * 16); tile0 < n; tile0 += stride) {
        // layer 0 operand: features s * 32 + g * 8 + {0..7} of sample nt * 16 + r
        bf16x8 xb[NT][2];
#pragma unroll
        for (int nt = 0; nt < NT; nt++)
#pragma unroll
            for (int s = 0; s < 2; s++) xb[nt][s] = x_next[nt][s];
        if (tile0 + stride < n) {
            load_x(rows_next, x_next);
#pragma unroll
            for (int nt = 0; nt < NT; nt++) rows_next[nt] = row_of(tile0 + 2 * stride, nt);
        }
#pragma unroll
        for (int ly = 0; ly < 4; ly++) {
            f32x4 acc[4][NT];
#pragma unroll
            for (int mt = 0; mt < 4; mt++)
#pragma unroll
                for (int nt = 0; nt < NT; nt++) acc[mt][nt] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int s = 0; s < 2; s++)
#pragma unroll
                for (int mt = 0; mt < 4; mt++)
#pragma unroll
                    for (int nt = 0; nt < NT; nt++) acc[mt][nt] = mfma(A[ly][mt][s], xb[nt][s], acc[mt][nt]);
#pragma unroll
            for (int nt = 0; nt < NT; nt++)
#pragma unroll
                for (int s = 0; s < 2; s++) xb[nt][s] = relu_pack(acc[2 * s][nt], acc[2 * s + 1][nt]);
        }
        // output layer: one 16-output tile (outputs 0..2 are the rgb prediction)
        f32x4 o[NT];
#pragma unroll
        for (int nt = 0; nt < NT; nt++) o[nt] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int s = 0; s < 2; s++)
#pragma unroll
            for (int nt = 0; nt < NT; nt++) o[nt] = mfma(A4[s], xb[nt][s], o[nt]);
        if (g == 0) // lanes 0..15 hold outputs 0..3 of sample nt * 16 + r
#pragma unroll
            for (int nt = 0; nt < NT; nt++) {
                const int i = tile0 + nt * 16 + r;
                if (i < n) out[i] = make_float4(o[nt][0], o[nt][1], o[nt][2], 0.f);
            }
    }
}
#ifndef TPT_NRC_INFER_TILES
#define TPT_NRC_INFER_TILES 2
#endif
static constexpr int kInferTiles = TPT_NRC_INFER_TILES; // 16-record tiles per wavefront and step
// resident inference wavefronts per CU (two per SIMD: the weights take 136 registers)
#ifndef TPT_NRC_INFER_WAVES_PER_CU
#define TPT_NRC_INFER_WAVES_PER_CU 8
#endif

// ---------------------------------------------------------------- training
// Forward + backward over 64-sample tiles on workgroups of four wavefronts (a
// tile's five activation matrices and two gradient matrices fill ~64 KB of
// LDS).  A tile's 4 x 4 output blocks of 16 x 16 are split over the waves: in
// the forward pass and the input gradient wave w computes feature block w of
// every sample block, in the weight gradient input-feature block w of every
// output block.  Every sum runs over K in the same groups of 8 as with one
// wavefront.  Loss: relative L2, sum_c (p_c - t_c)^2 / (sg(p_c)^2 + 0.01),
// averaged over the n samples; the mean loss is added to *loss.
// A workgroup walks tiles blockIdx.x, blockIdx.x + gridDim.x, .. and sums its
// tiles' weight gradients into
// partial[blockIdx.x] (W[l][in][out] layout; the output layer's padding columns
// are not written, zero without a tile), and the Adam kernel sums the gridDim.x
// partials in a fixed order.  (One FP32 atomic
// per weight and tile into one gradient, 2.1M atomics per step of 8192 records,
// took 38 us per step alone; the partials move 10 MB instead.)
__device__ __forceinline__ uint16_t bf16_one(float f) { return (uint16_t)bf16_pair(f, 0.f); }
static constexpr int kTrainWaves = 4;
__global__ __launch_bounds__(64 * kTrainWaves) void k_nrc_train(
    const uint16_t *__restrict__ feats, const int32_t *__restrict__ idx, const float4 *__restrict__ target,
    const uint32_t *__restrict__ n_dev, int cap, int base, int n_max, const uint16_t *__restrict__ wt,
    const uint16_t *__restrict__ wb, float *__restrict__ partial, float *__restrict__ loss) {
    __shared__ __attribute__((aligned(16))) uint16_t act[5][64 * kRow]; // layer inputs x0..x4
    __shared__ __attribute__((aligned(16))) uint16_t g[2][64 * kRow];   // output gradients (ping-pong)
    // records base .. base + n - 1 of the first min(*n_dev, cap): the producer keeps
    // counting candidates past the `cap` records it stores
    const int n = min(max(min((int)*n_dev, cap) - base, 0), n_max);
    if (idx) idx += base;
    target += base;
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    const int r = l & 15, q = l >> 4;
    // this workgroup's gradient: written by its first tile, added to by the others
    float *dst = partial + (size_t)blockIdx.x * kNrcParams;
    const float inv_n = n > 0 ? 1.f / (float)n : 0.f;
    float lsum = 0.f;
    for (int tile0 = blockIdx.x * 64; tile0 < n; tile0 += gridDim.x * 64) { // workgroup-uniform
        // the tile's feature rows (zeros beyond n), 16 B per thread and step
        for (int i = threadIdx.x; i < 64 * 8; i += 64 * kTrainWaves) {
            const int s = i >> 3, c = (i & 7) * 8;
            uint4 v = make_uint4(0, 0, 0, 0);
            if (tile0 + s < n) v = *reinterpret_cast<const uint4 *>(rec_row(feats, idx, tile0 + s) + c);
            *reinterpret_cast<uint4 *>(act[0] + s * kRow + c) = v;
        }
        __syncthreads();
        // hidden layers: wave w computes outputs w * 16 .. w * 16 + 15 of all 64 samples
#pragma nounroll
        for (int ly = 0; ly < 4; ly++) {
            const uint16_t *x = act[ly];
            const uint16_t *W = wt + ly * 4096;
            f32x4 acc[4];
#pragma unroll
            for (int mt = 0; mt < 4; mt++) acc[mt] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int ks = 0; ks < 2; ks++) {
                const bf16x8 b = ld8(W + (w * 16 + r) * 64 + ks * 32 + q * 8);
#pragma unroll
                for (int mt = 0; mt < 4; mt++)
                    acc[mt] = mfma(ld8_lds(x + (mt * 16 + r) * kRow + ks * 32 + q * 8), b, acc[mt]);
            }
            uint16_t *y = act[ly + 1];
#pragma unroll
            for (int mt = 0; mt < 4; mt++)
#pragma unroll
                for (int j = 0; j < 4; j++)
                    y[(mt * 16 + q * 4 + j) * kRow + w * 16 + r] = bf16_one(fmaxf(acc[mt][j], 0.f));
            __syncthreads();
        }
        // output layer (outputs 0..2 in block 0): wave w computes samples w * 16 .. w * 16 + 15
        f32x4 pred = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int ks = 0; ks < 2; ks++)
            pred = mfma(ld8_lds(act[4] + (w * 16 + r) * kRow + ks * 32 + q * 8),
                        ld8(wt + 4 * 4096 + r * 64 + ks * 32 + q * 8), pred);
        // loss gradient of the output (columns 0..2), written as BF16 into g[0][s][0..31]
        // (columns 3..31 zero: the last layer's K step for the backward pass is 32)
        for (int i = threadIdx.x; i < 64 * 32; i += 64 * kTrainWaves) g[0][(i >> 5) * kRow + (i & 31)] = 0;
        __syncthreads();
        if (r < 3) {
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const int s = w * 16 + q * 4 + j;
                float gp = 0.f;
                if (tile0 + s < n) {
                    const float4 t4 = target[tile0 + s];
                    const float t = r == 0 ? t4.x : (r == 1 ? t4.y : t4.z);
                    const float p = pred[j];
                    const float den = p * p + 0.01f;
                    const float d = p - t;
                    lsum += d * d / den * inv_n;
                    gp = 2.f * d / den * inv_n;
                }
                g[0][s * kRow + r] = bf16_one(gp);
            }
        }
        __syncthreads();
        // backward, last layer first: go = g[cur] is dL/d(output of layer ly)
        int cur = 0;
#pragma nounroll
        for (int ly = 4; ly >= 0; ly--) {
            const uint16_t *x = act[ly];
            const uint16_t *go = g[cur];
            const int NO = ly == 4 ? 1 : 4; // output blocks of 16 (the last layer's 3 outputs sit in block 0)
            const int KO = ly == 4 ? 1 : 2; // K steps of 32 over the outputs
            // dW[in][out] = sum_s x[s][in] * go[s][out]  (M = in: block w, N = out, K = samples)
            {
                f32x4 dw[4];
#pragma unroll
                for (int nt = 0; nt < 4; nt++) dw[nt] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int ks = 0; ks < 2; ks++) {
                    const bf16x8 a = gather8(x + (ks * 32 + q * 8) * kRow + w * 16 + r, kRow);
#pragma unroll
                    for (int nt = 0; nt < NO; nt++)
                        dw[nt] = mfma(a, gather8(go + (ks * 32 + q * 8) * kRow + nt * 16 + r, kRow), dw[nt]);
                }
                const bool first = tile0 == (int)blockIdx.x * 64;
#pragma unroll
                for (int nt = 0; nt < NO; nt++)
#pragma unroll
                    for (int j = 0; j < 4; j++) {
                        const int in = w * 16 + q * 4 + j, out = nt * 16 + r;
                        if (ly == 4 && out >= 3) continue;
                        float *d = dst + ly * 4096 + in * 64 + out;
                        *d = first ? dw[nt][j] : *d + dw[nt][j];
                    }
            }
            if (ly == 0) break;
            // gin[s][in] = (sum_o go[s][o] * W[in][o]) * (x[s][in] > 0)  (M = samples, N = in: block w, K = out)
            uint16_t *gi = g[cur ^ 1];
            const uint16_t *wl = wb + ly * 4096;
            f32x4 acc[4];
#pragma unroll
            for (int mt = 0; mt < 4; mt++) acc[mt] = f32x4{0.f, 0.f, 0.f, 0.f};
            for (int ks = 0; ks < KO; ks++) {
                const bf16x8 b = ld8(wl + (w * 16 + r) * 64 + ks * 32 + q * 8);
#pragma unroll
                for (int mt = 0; mt < 4; mt++)
                    acc[mt] = mfma(ld8_lds(go + (mt * 16 + r) * kRow + ks * 32 + q * 8), b, acc[mt]);
            }
#pragma unroll
            for (int mt = 0; mt < 4; mt++)
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    const int s = mt * 16 + q * 4 + j, in = w * 16 + r;
                    const float relu_on = bf2f(x[s * kRow + in]) > 0.f ? 1.f : 0.f;
                    gi[s * kRow + in] = bf16_one(acc[mt][j] * relu_on);
                }
            __syncthreads();
            cur ^= 1;
        }
        __syncthreads(); // the next tile overwrites act and g
    }
    for (int off = 32; off > 0; off >>= 1) lsum += __shfl_down(lsum, off);
    if (l == 0 && lsum != 0.f) atomicAdd(loss, lsum);
    if (blockIdx.x * 64 >= n) // no tile: a zero gradient
        for (int i = threadIdx.x; i < kNrcParams; i += 64 * kTrainWaves) dst[i] = 0.f;
}

// Adam (Kingma & Ba 2015) on the FP32 master weights, then the two BF16 copies.
// The gradient is the sum of the training kernel's `groups` partials, in a fixed
// order: a workgroup of kAdamSplit x 64 threads takes 64 consecutive weights,
// wavefront r sums partials r, r + kAdamSplit, .. (64 consecutive floats per
// load), and the kAdamSplit sums are added in LDS.  The sum is also stored in
// grad (read by the test hooks).  t: step number (1-based).
static constexpr int kAdamSplit = 4;
__global__ __launch_bounds__(64 * kAdamSplit) void k_nrc_adam(float *__restrict__ w, float *__restrict__ m,
                                                               float *__restrict__ v, const float *__restrict__ partial,
                                                               int groups, float *__restrict__ grad,
                                                               uint16_t *__restrict__ wt, uint16_t *__restrict__ wb,
                                                               float lr, float b1, float b2, float eps, float bc1,
                                                               float bc2) {
    __shared__ float part[kAdamSplit][64];
    const int lane = threadIdx.x & 63, r = threadIdx.x >> 6;
    const int i = blockIdx.x * 64 + lane; // 5 * 4096 is a multiple of 64
    float gr = 0.f;
#pragma unroll 8
    for (int k = r; k < groups; k += kAdamSplit) gr += partial[(size_t)k * kNrcParams + i];
    part[r][lane] = gr;
    __syncthreads();
    if (r != 0) return;
    gr = 0.f;
#pragma unroll
    for (int k = 0; k < kAdamSplit; k++) gr += part[k][lane];
    int ly = i >> 12, in = (i >> 6) & 63, out = i & 63;
    if (ly == 4 && out >= 3) gr = 0.f; // padding columns (not written by the training kernel)
    grad[i] = gr;
    float mi = b1 * m[i] + (1.f - b1) * gr;
    float vi = b2 * v[i] + (1.f - b2) * gr * gr;
    m[i] = mi;
    v[i] = vi;
    float wi = w[i];
    if (ly < 4 || out < 3) wi -= lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
    else wi = 0.f;
    w[i] = wi;
    uint16_t b = f2bf(wi);
    wb[i] = b;                                // [l][in][out]
    wt[(ly << 12) + out * 64 + in] = b;       // [l][out][in]
}

// ---------------------------------------------------------------- launchers
void nrc_launch_infer(const uint16_t *feats, const int32_t *idx, const uint32_t *n_dev, int n_max,
                      const uint16_t *wt, float4 *out, int num_cu, hipStream_t st) {
    if (n_max <= 0) return;
    const int per_wave = kInferTiles * 16;
    const int waves = std::min((n_max + per_wave - 1) / per_wave, num_cu * TPT_NRC_INFER_WAVES_PER_CU);
    const int blocks = (waves + 3) / 4;
    hipLaunchKernelGGL(k_nrc_infer<kInferTiles>, dim3(blocks), dim3(256), 0, st, feats, idx, n_dev, n_max, wt, out);
}
int nrc_train_groups(int n_max) { return std::max(1, std::min((n_max + 63) / 64, kNrcMaxTrainGroups)); }
void nrc_launch_train(const uint16_t *feats, const int32_t *idx, const float4 *target, const uint32_t *n_dev,
                      int cap, int base, int n_max, const NrcWeights &W, hipStream_t st) {
    if (n_max <= 0) return;
    hipLaunchKernelGGL(k_nrc_train, dim3(nrc_train_groups(n_max)), dim3(64 * kTrainWaves), 0, st, feats, idx, target,
                       n_dev, cap, base, n_max, W.wt, W.wb, W.partial, W.loss);
}
void nrc_launch_adam(const NrcWeights &W, int n_max, float lr, int step, hipStream_t st) {
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    float bc1 = 1.f - powf(b1, (float)step), bc2 = 1.f - powf(b2, (float)step);
    hipLaunchKernelGGL(k_nrc_adam, dim3(kNrcParams / 64), dim3(64 * kAdamSplit), 0, st, W.w, W.m, W.v, W.partial,
                       nrc_train_groups(n_max), W.grad, W.wt, W.wb, lr, b1, b2, eps, bc1, bc2);
}

} // namespace tpt
