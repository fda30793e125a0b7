// mlp_device.h -- the warm-start MLP on the device: the ray encoding and the
// BF16 MFMA forward pass of one wave (64 rays, one per lane).
//
// Network (no counterpart in the reference): x = (origin / 4, unit
// direction); features sin(2^k pi x_i), cos(2^k pi x_i) for k = 0..5 (72),
// then x (6), padded to 80; 80 -> 64 -> 64 -> 64 -> 2, ReLU; outputs
// (hit distance / 4, hit logit).
//
// MFMA: v_mfma_f32_16x16x32_bf16 (gfx950, wave64), and its K = 16 form for
// the last 16 inputs of the first layer.  Lane l = 16 q + m:
//   A (16 x K, M x K):  row m, 8 (K = 32) or 4 (K = 16) values of group q
//   B (K x 16, K x N):  column m, the same K slots of group q
//   D (16 x 16, M x N): D[4 q + i][m], i = 0..3
// The MFMA sums over the K slots, so any assignment of inputs to the
// (q, element) slots works if the weights (A) and activations (B) use the
// same one.  Each layer runs transposed (out^T = W^T in^T: M = output
// features, N = rays), so that D is the next layer's B: a K = 32 tile of a
// hidden layer holds, in lane group q, inputs 32 kt + 4 q + i (elements
// 0..3, from the previous layer's D tile 2 kt) and 32 kt + 16 + 4 q + i
// (elements 4..7, D tile 2 kt + 1).  The first layer's K = 32 tiles take
// features 32 kt + 8 q + i from LDS, its K = 16 tile features 64 + 4 q + i.
// The weights are packed in that order per tile (16 or 8 bytes per lane)
// and staged in LDS once per block; biases stay f32.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/mb3d_hip.h"

namespace mb3d {

constexpr int MLP_IN = 6, MLP_FREQ = 6, MLP_K1 = 80, MLP_H = 64, MLP_OUT = 2;
constexpr int MLP_PARAMS = MLP_K1 * MLP_H + MLP_H + 2 * (MLP_H * MLP_H + MLP_H) + MLP_H * MLP_OUT + MLP_OUT;
// packed bf16 weights (u16 offsets): layer 1's 2 x 4 K = 32 tiles and 4 K = 16
// tiles, layers 2 and 3 2 x 4 K = 32 tiles each, layer 4 2 K = 32 tiles
constexpr int MLP_T32 = 64 * 8, MLP_T16 = 64 * 4;
constexpr int MLP_OFF_L1A = 0, MLP_OFF_L1B = 8 * MLP_T32, MLP_OFF_L2 = MLP_OFF_L1B + 4 * MLP_T16;
constexpr int MLP_OFF_L3 = MLP_OFF_L2 + 8 * MLP_T32, MLP_OFF_L4 = MLP_OFF_L3 + 8 * MLP_T32;
constexpr int MLP_PACKED_U16 = MLP_OFF_L4 + 2 * MLP_T32;
constexpr int MLP_TILES = MLP_PACKED_U16 / (64 * 4);  // in 64-lane units of 4 values
constexpr int MLP_BIAS = 3 * MLP_H + 16;
// FLOPs of one ray's forward pass (multiply-adds counted twice)
constexpr int MLP_FLOP = 2 * (MLP_K1 * MLP_H + 2 * MLP_H * MLP_H + MLP_H * MLP_OUT);

// the network as the march kernels take it
struct MlpDev {
    const uint16_t *w;  // MLP_PACKED_U16 bf16 values, tile-major
    const float *b;     // MLP_BIAS: b1, b2, b3, b4 (padded to 16)
    float conf_hit;     // sigmoid(logit) above: a jump to the predicted hit
    float conf_miss;    // at most: background without a march (< 0: never)
    float margin_rel;   // the jump stops short of the prediction by this share
    float margin_abs;   // ... and this many scene units
    int enabled;
    // the cone pass's gate: a tile whose middle ray's hit logit is at least
    // gate_logit marches no cone ray (gate 0: off)
    int gate;
    float gate_logit;
};

// (host) the device form of `net` with the options' thresholds; disabled
// (enabled = 0) when net is NULL
void mlp_dev(const mb3d_hip_mlp *net, const mb3d_hip_options &o, MlpDev &M);

typedef short mlp_s4 __attribute__((ext_vector_type(4)));
typedef short mlp_s8 __attribute__((ext_vector_type(8)));
typedef __bf16 mlp_b8 __attribute__((ext_vector_type(8)));
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

// ReLU of an accumulator tile as a half of the next layer's B fragment
__device__ __forceinline__ mlp_s4 mlp_relu_frag(mlp_f4 acc) {
    const uint32_t lo = mlp_pk(fmaxf(acc[0], 0.0f), fmaxf(acc[1], 0.0f));
    const uint32_t hi = mlp_pk(fmaxf(acc[2], 0.0f), fmaxf(acc[3], 0.0f));
    return mlp_s4{(short)(lo & 0xFFFF), (short)(lo >> 16), (short)(hi & 0xFFFF), (short)(hi >> 16)};
}

__device__ __forceinline__ mlp_f4 mlp_mfma32(mlp_s8 a, mlp_s8 b, mlp_f4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(mlp_b8, a), __builtin_bit_cast(mlp_b8, b), c, 0,
                                                  0, 0);
}

// two D tiles of a layer (features 32 kt + 4 q + i, then 32 kt + 16 + 4 q +
// i) as one K = 32 B fragment
__device__ __forceinline__ mlp_s8 mlp_cat(mlp_s4 a, mlp_s4 b) {
    return __builtin_shufflevector(a, b, 0, 1, 2, 3, 4, 5, 6, 7);
}

// A hidden layer (64 -> 64), transposed, for the wave's 4 groups of 16
// rays: each weight tile (read once from LDS) serves the 4 groups, and the
// accumulators (features x rays), which start at the bias, are the next
// layer's B fragments -- no LDS round trip between layers.
__device__ __forceinline__ void mlp_hidden(const mlp_s4 (&in)[4][4], const mlp_s8 *w, const float *__restrict__ bias,
                                           mlp_s4 (&out)[4][4]) {
    const int l = __lane_id();
#pragma unroll
    for (int nt = 0; nt < 4; nt++) {
        const float4 b4 = *reinterpret_cast<const float4 *>(bias + 16 * nt + 4 * (l >> 4));
        mlp_f4 acc[4];
#pragma unroll
        for (int g = 0; g < 4; g++) acc[g] = mlp_f4{b4.x, b4.y, b4.z, b4.w};
#pragma unroll
        for (int kt = 0; kt < 2; kt++) {
            const mlp_s8 a = w[(kt * 4 + nt) * 64 + l];
#pragma unroll
            for (int g = 0; g < 4; g++) acc[g] = mlp_mfma32(a, mlp_cat(in[g][2 * kt], in[g][2 * kt + 1]), acc[g]);
        }
#pragma unroll
        for (int g = 0; g < 4; g++) out[g][nt] = mlp_relu_frag(acc[g]);
    }
}

// The forward pass for the 64 rays of a wave (lane i's ray: origin o,
// unit direction d): returns the lane's (hit distance / 4, hit logit).  All
// 64 lanes must be active.  `lds` is this wave's MLP_LDS_BYTES, `wl` the
// block's staged weights (16-byte aligned).
__device__ __forceinline__ void mlp_wave_infer(const MlpDev &M, const mlp_s4 *wl, const float o[3], const float d[3],
                                               char *lds, float &t_out, float &logit_out) {
    uint16_t *feat = reinterpret_cast<uint16_t *>(lds);
    const int l = __lane_id(), q = l >> 4;
    // the lane's own ray's features, then the B fragments of all 4 groups
    // (group g: rays 16 g .. 16 g + 15, lane l% 16's ray)
    mlp_features(o, d, feat + l * MLP_K1);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    mlp_s8 x32[4][2];
    mlp_s4 x16[4];
#pragma unroll
    for (int g = 0; g < 4; g++) {
        const uint16_t *row = feat + (16 * g + (l & 15)) * MLP_K1;
#pragma unroll
        for (int kt = 0; kt < 2; kt++) x32[g][kt] = *reinterpret_cast<const mlp_s8 *>(row + 32 * kt + 8 * q);
        x16[g] = *reinterpret_cast<const mlp_s4 *>(row + 64 + 4 * q);
    }
    __builtin_amdgcn_wave_barrier();
    const uint16_t *wb = reinterpret_cast<const uint16_t *>(wl);
    const mlp_s8 *w1a = reinterpret_cast<const mlp_s8 *>(wb + MLP_OFF_L1A);
    const mlp_s4 *w1b = reinterpret_cast<const mlp_s4 *>(wb + MLP_OFF_L1B);
    const mlp_s8 *w2 = reinterpret_cast<const mlp_s8 *>(wb + MLP_OFF_L2);
    const mlp_s8 *w3 = reinterpret_cast<const mlp_s8 *>(wb + MLP_OFF_L3);
    const mlp_s8 *w4 = reinterpret_cast<const mlp_s8 *>(wb + MLP_OFF_L4);
    mlp_s4 h1[4][4], h2[4][4];
    // layer 1: 80 -> 64 (two K = 32 tiles and one K = 16 tile)
#pragma unroll
    for (int nt = 0; nt < 4; nt++) {
        const float4 b4 = *reinterpret_cast<const float4 *>(M.b + 16 * nt + 4 * q);
        mlp_f4 acc[4];
#pragma unroll
        for (int g = 0; g < 4; g++) acc[g] = mlp_f4{b4.x, b4.y, b4.z, b4.w};
#pragma unroll
        for (int kt = 0; kt < 2; kt++) {
            const mlp_s8 a = w1a[(kt * 4 + nt) * 64 + l];
#pragma unroll
            for (int g = 0; g < 4; g++) acc[g] = mlp_mfma32(a, x32[g][kt], acc[g]);
        }
        const mlp_s4 a16 = w1b[nt * 64 + l];
#pragma unroll
        for (int g = 0; g < 4; g++) acc[g] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a16, x16[g], acc[g], 0, 0, 0);
#pragma unroll
        for (int g = 0; g < 4; g++) h1[g][nt] = mlp_relu_frag(acc[g]);
    }
    mlp_hidden(h1, w2, M.b + MLP_H, h2);
    mlp_hidden(h2, w3, M.b + 2 * MLP_H, h1);
    // the output layer (N = 2 of a 16-wide tile): lanes 0..15 hold rows 0..3
    // (distance, logit, 0, 0) of ray 16 g + lane
    const float4 b4 = *reinterpret_cast<const float4 *>(M.b + 3 * MLP_H + 4 * q);
    float t = 0.0f, lg = 0.0f;
#pragma unroll
    for (int g = 0; g < 4; g++) {
        mlp_f4 acc = {b4.x, b4.y, b4.z, b4.w};
#pragma unroll
        for (int kt = 0; kt < 2; kt++) acc = mlp_mfma32(w4[kt * 64 + l], mlp_cat(h1[g][2 * kt], h1[g][2 * kt + 1]), acc);
        const float tg = __shfl(acc[0], l & 15, 64), lgg = __shfl(acc[1], l & 15, 64);
        if (q == g) {
            t = tg;
            lg = lgg;
        }
    }
    t_out = t;
    logit_out = lg;
}

}  // namespace mb3d
