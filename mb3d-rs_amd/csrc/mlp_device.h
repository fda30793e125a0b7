// mlp_device.h -- the warm-start MLP on the device: the ray encoding and the
// BF16 MFMA forward pass of one wave (64 rays, one per lane).
//
// Network (no counterpart in the reference): x = (origin / 4, unit
// direction); features sin(2^k pi x_i), cos(2^k pi x_i) for k = 0..5 (72),
// then x (6), padded to 80; 80 -> 64 -> 64 -> 64 -> 2, ReLU; outputs
// (hit distance / 4, hit logit).
//
// MFMA: v_mfma_f32_16x16x16_bf16 (wave64).  Operand layout, lane l:
//   A (16 x 16, M x K):   A[l % 16][4 (l / 16) + i], i = 0..3
//   B (16 x 16, K x N):   B[4 (l / 16) + i][l % 16]
//   D (16 x 16, M x N):   D[4 (l / 16) + i][l % 16]
// Each layer runs transposed (out^T = W^T in^T: M = output features, N =
// rays), so that D is the next layer's B.  The weights are packed per tile
// as the A operand W^T (lane l: W[16 kt + 4 (l / 16) + i][16 nt + l % 16],
// 8 bytes per lane per tile) and staged in LDS once per block; biases stay
// f32.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/mb3d_hip.h"

namespace mb3d {

constexpr int MLP_IN = 6, MLP_FREQ = 6, MLP_K1 = 80, MLP_H = 64, MLP_OUT = 2;
constexpr int MLP_PARAMS = MLP_K1 * MLP_H + MLP_H + 2 * (MLP_H * MLP_H + MLP_H) + MLP_H * MLP_OUT + MLP_OUT;
// packed bf16 tiles: layer 1 5 x 4, layers 2 and 3 4 x 4, layer 4 4 x 1
constexpr int MLP_TILES = 5 * 4 + 4 * 4 + 4 * 4 + 4 * 1;
constexpr int MLP_PACKED_U16 = MLP_TILES * 64 * 4;
constexpr int MLP_BIAS = 3 * MLP_H + 16;

// the network as the march kernels take it
struct MlpDev {
    const uint16_t *w;  // MLP_PACKED_U16 bf16 values, tile-major
    const float *b;     // MLP_BIAS: b1, b2, b3, b4 (padded to 16)
    float conf_hit;     // sigmoid(logit) above: a jump to the predicted hit
    float conf_miss;    // at most: background without a march (< 0: never)
    float margin_rel;   // the jump stops short of the prediction by this share
    float margin_abs;   // ... and this many scene units
    int enabled;
};

// (host) the device form of `net` with the options' thresholds; disabled
// (enabled = 0) when net is NULL
void mlp_dev(const mb3d_hip_mlp *net, const mb3d_hip_options &o, MlpDev &M);

typedef short mlp_s4 __attribute__((ext_vector_type(4)));
typedef float mlp_f4 __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(1))) const mlp_s4 mlp_gs4;  // a weight tile's lane share, global memory

typedef __bf16 mlp_bf2 __attribute__((ext_vector_type(2)));
typedef float mlp_f2 __attribute__((ext_vector_type(2)));

// (a, b) as a bf16 pair, round to nearest even (v_cvt_pk_bf16_f32)
__device__ __forceinline__ uint32_t mlp_pk(float a, float b) {
    const mlp_f2 v = {a, b};
    return __builtin_bit_cast(uint32_t, __builtin_convertvector(v, mlp_bf2));
}

// the 80 features of a ray (origin, unit direction), written as bf16 to
// `row` (16-byte aligned)
__device__ __forceinline__ void mlp_features(const float o[3], const float d[3], uint16_t *row) {
    const float x[6] = {o[0] * 0.25f, o[1] * 0.25f, o[2] * 0.25f, d[0], d[1], d[2]};
    uint32_t w[40];
    int c = 0;
#pragma unroll
    for (int i = 0; i < MLP_IN; i++) {
        // sin and cos of pi x (the hardware's v_sin/v_cos take revolutions;
        // |x| <= ~1 here), then the double angle for each octave
        float s = __builtin_amdgcn_sinf(0.5f * x[i]), co = __builtin_amdgcn_cosf(0.5f * x[i]);
#pragma unroll
        for (int k = 0; k < MLP_FREQ; k++) {
            w[c++] = mlp_pk(s, co);
            const float s2 = 2.0f * s * co, c2 = 1.0f - 2.0f * s * s;
            s = s2;
            co = c2;
        }
    }
    w[36] = mlp_pk(x[0], x[1]);
    w[37] = mlp_pk(x[2], x[3]);
    w[38] = mlp_pk(x[4], x[5]);
    w[39] = 0u;
    uint4 *dst = reinterpret_cast<uint4 *>(row);
#pragma unroll
    for (int j = 0; j < 10; j++) dst[j] = make_uint4(w[4 * j], w[4 * j + 1], w[4 * j + 2], w[4 * j + 3]);
}

// LDS of one wave for mlp_wave_infer (bytes): the 64 rays' feature rows
constexpr int MLP_LDS_BYTES = 64 * MLP_K1 * 2;
// the packed weights, staged in LDS once per block (bytes)
constexpr int MLP_W_LDS_BYTES = MLP_PACKED_U16 * 2;

// Copies the packed weights to `wl` (LDS); every thread of the block calls
// it, and the block synchronises before the first mlp_wave_infer.
__device__ __forceinline__ void mlp_stage_weights(const MlpDev &M, mlp_s4 *wl) {
    const mlp_gs4 *src = (const mlp_gs4 *)M.w;
    for (int i = threadIdx.x; i < MLP_TILES * 64; i += blockDim.x) wl[i] = src[i];
}

// ReLU of an accumulator tile as the next layer's B fragment
__device__ __forceinline__ mlp_s4 mlp_relu_frag(mlp_f4 acc) {
    const uint32_t lo = mlp_pk(fmaxf(acc[0], 0.0f), fmaxf(acc[1], 0.0f));
    const uint32_t hi = mlp_pk(fmaxf(acc[2], 0.0f), fmaxf(acc[3], 0.0f));
    return mlp_s4{(short)(lo & 0xFFFF), (short)(lo >> 16), (short)(hi & 0xFFFF), (short)(hi >> 16)};
}

// One layer, transposed (out^T = W^T in^T), for the wave's 4 groups of 16
// rays: the weight tile is the A operand (read once, used by all 4 groups),
// the activations the B operand, and the accumulator D (features x rays) is
// already the next layer's B fragment -- no LDS round trip between layers.
// The accumulators start at the bias.
template <int KT, int NT>
__device__ __forceinline__ void mlp_layer(const mlp_s4 (&in)[4][KT], const mlp_s4 *w, const float *__restrict__ bias,
                                          mlp_s4 (&out)[4][NT]) {
    const int l = __lane_id();
#pragma unroll
    for (int nt = 0; nt < NT; nt++) {
        const float4 b4 = *reinterpret_cast<const float4 *>(bias + 16 * nt + 4 * (l >> 4));
        mlp_f4 acc[4];
#pragma unroll
        for (int g = 0; g < 4; g++) acc[g] = mlp_f4{b4.x, b4.y, b4.z, b4.w};
#pragma unroll
        for (int kt = 0; kt < KT; kt++) {
            const mlp_s4 a = w[(kt * NT + nt) * 64 + l];
#pragma unroll
            for (int g = 0; g < 4; g++) acc[g] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, in[g][kt], acc[g], 0, 0, 0);
        }
#pragma unroll
        for (int g = 0; g < 4; g++) out[g][nt] = mlp_relu_frag(acc[g]);
    }
}

// The forward pass for the 64 rays of a wave (lane i's ray: origin o,
// unit direction d): returns the lane's (hit distance / 4, hit logit).  All
// 64 lanes must be active.  `lds` is this wave's MLP_LDS_BYTES, `wl` the
// block's staged weights.
__device__ __forceinline__ void mlp_wave_infer(const MlpDev &M, const mlp_s4 *wl, const float o[3], const float d[3],
                                               char *lds, float &t_out, float &logit_out) {
    uint16_t *feat = reinterpret_cast<uint16_t *>(lds);
    const int l = __lane_id();
    // the lane's own ray's features, then the B fragments of all 4 groups:
    // group g, k-tile kt holds features 16 kt + 4 (l / 16) .. + 3 of ray
    // 16 g + l % 16
    mlp_features(o, d, feat + l * MLP_K1);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    mlp_s4 x0[4][5];
#pragma unroll
    for (int g = 0; g < 4; g++)
#pragma unroll
        for (int kt = 0; kt < 5; kt++)
            x0[g][kt] = *reinterpret_cast<const mlp_s4 *>(feat + (16 * g + (l & 15)) * MLP_K1 + 16 * kt + 4 * (l >> 4));
    __builtin_amdgcn_wave_barrier();
    const mlp_s4 *w1 = wl;
    const mlp_s4 *w2 = w1 + 5 * 4 * 64;
    const mlp_s4 *w3 = w2 + 4 * 4 * 64;
    const mlp_s4 *w4 = w3 + 4 * 4 * 64;
    mlp_s4 h1[4][4], h2[4][4];
    mlp_layer<5, 4>(x0, w1, M.b, h1);
    mlp_layer<4, 4>(h1, w2, M.b + MLP_H, h2);
    mlp_layer<4, 4>(h2, w3, M.b + 2 * MLP_H, h1);
    // the output layer (N = 2 of a 16-wide tile): lanes 0..15 hold rows 0..3
    // (distance, logit, 0, 0) of ray 16 g + lane
    const float4 b4 = *reinterpret_cast<const float4 *>(M.b + 3 * MLP_H + 4 * (l >> 4));
    float t = 0.0f, lg = 0.0f;
#pragma unroll
    for (int g = 0; g < 4; g++) {
        mlp_f4 acc = {b4.x, b4.y, b4.z, b4.w};
#pragma unroll
        for (int kt = 0; kt < 4; kt++) acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(w4[kt * 64 + l], h1[g][kt], acc, 0, 0, 0);
        const float tg = __shfl(acc[0], l & 15, 64), lgg = __shfl(acc[1], l & 15, 64);
        if ((l >> 4) == g) {
            t = tg;
            lg = lgg;
        }
    }
    t_out = t;
    logit_out = lg;
}

}  // namespace mb3d
