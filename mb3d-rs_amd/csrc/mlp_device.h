// mlp_device.h -- the warm-start MLP on the device: the ray encoding and the
// BF16 MFMA forward pass of one wave (64 rays, one per lane).
//
// Network (no counterpart in the reference): x = (origin / 4, unit
// direction); features sin(2^k pi x_i), cos(2^k pi x_i) for k = 0..5 (72),
// then x (6), padded to 80; 80 -> 64 -> 64 -> 64 -> 2, ReLU; outputs
// (hit distance / 4, hit logit).
//
// MFMA: v_mfma_f32_16x16x16_bf16 (wave64).  Operand layout, lane l:
//   A (16 x 16, rows = rays):   A[l % 16][4 (l / 16) + i], i = 0..3
//   B (16 x 16, K x N):         B[4 (l / 16) + i][l % 16]
//   D (16 x 16):                D[4 (l / 16) + i][l % 16]
// The weights are packed per tile in exactly the B layout (8 bytes per lane
// per tile: one coalesced load), biases stay f32.
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
    float conf_miss;    // below: a jump to the predicted free distance
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

__device__ __forceinline__ uint16_t f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even
    return (uint16_t)(u >> 16);
}

// the 80 features of a ray (origin, unit direction), written as bf16 to
// `row`
__device__ __forceinline__ void mlp_features(const float o[3], const float d[3], uint16_t *row) {
    const float x[6] = {o[0] * 0.25f, o[1] * 0.25f, o[2] * 0.25f, d[0], d[1], d[2]};
    int c = 0;
#pragma unroll
    for (int i = 0; i < MLP_IN; i++) {
        // sin and cos of pi x (the hardware's v_sin/v_cos take revolutions;
        // |x| <= ~1 here), then the double angle for each octave
        float s = __builtin_amdgcn_sinf(0.5f * x[i]), co = __builtin_amdgcn_cosf(0.5f * x[i]);
#pragma unroll
        for (int k = 0; k < MLP_FREQ; k++) {
            row[c] = f2bf(s);
            row[c + 1] = f2bf(co);
            c += 2;
            const float s2 = 2.0f * s * co, c2 = 1.0f - 2.0f * s * s;
            s = s2;
            co = c2;
        }
    }
#pragma unroll
    for (int i = 0; i < MLP_IN; i++) row[c + i] = f2bf(x[i]);
    row[78] = 0;
    row[79] = 0;
}

// LDS of one wave for mlp_wave_infer (bytes)
constexpr int MLP_LDS_BYTES = 16 * MLP_K1 * 2 + 16 * MLP_H * 2 + 16 * 2 * 4;

// One MFMA layer for the 16 rays of an m-tile: A fragments from `in`
// (row stride `ld` bf16), K = 16 * kt_n, N = 16 * nt_n; the result (bias,
// ReLU) into `out` as bf16 (row stride MLP_H), or the 2 outputs as f32.
template <int KT, int NT, bool LAST>
__device__ __forceinline__ void mlp_layer(const uint16_t *in, int ld, const uint16_t *w,
                                          const float *__restrict__ bias, uint16_t *out, float *outf) {
    const int l = __lane_id();
    const int row = l & 15, kq = (l >> 4) * 4;
    mlp_s4 a[KT];
#pragma unroll
    for (int kt = 0; kt < KT; kt++) {
        const uint16_t *p = in + row * ld + kt * 16 + kq;
        a[kt] = mlp_s4{(short)p[0], (short)p[1], (short)p[2], (short)p[3]};
    }
    __builtin_amdgcn_wave_barrier();
    // the weight tiles are loaded where they are used: an empty asm makes
    // the offset opaque, so that the compiler cannot hoist the loads of all
    // four layers out of the callers' loops and keep ~110 registers of
    // weights live through the march
    int woff = 0;
    asm volatile("" : "+s"(woff));
    const mlp_gs4 *wt = (const mlp_gs4 *)(w + woff);
#pragma unroll
    for (int nt = 0; nt < NT; nt++) {
        mlp_f4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int kt = 0; kt < KT; kt++) {
            const mlp_s4 b = wt[(kt * NT + nt) * 64 + l];
            acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[kt], b, acc, 0, 0, 0);
        }
        const int n = nt * 16 + (l & 15);
        const float bn = bias[n];
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int m = (l >> 4) * 4 + i;
            const float v = acc[i] + bn;
            if (LAST) {
                if (n < MLP_OUT) outf[m * 2 + n] = v;
            } else {
                out[m * MLP_H + n] = f2bf(fmaxf(v, 0.0f));
            }
        }
    }
    __builtin_amdgcn_wave_barrier();
}

// The forward pass for the 64 rays of a wave (lane i's ray: origin o,
// unit direction d): returns the lane's (hit distance / 4, hit logit).  All
// 64 lanes must be active.  `lds` is this wave's MLP_LDS_BYTES.
__device__ __forceinline__ void mlp_wave_infer(const MlpDev &M, const float o[3], const float d[3], char *lds,
                                               float &t_out, float &logit_out) {
    uint16_t *in = reinterpret_cast<uint16_t *>(lds);
    uint16_t *hid = in + 16 * MLP_K1;
    float *outf = reinterpret_cast<float *>(hid + 16 * MLP_H);
    const int l = __lane_id();
    const uint16_t *w1 = M.w;
    const uint16_t *w2 = w1 + 5 * 4 * 256;
    const uint16_t *w3 = w2 + 4 * 4 * 256;
    const uint16_t *w4 = w3 + 4 * 4 * 256;
#pragma unroll 1
    for (int mt = 0; mt < 4; mt++) {
        if ((l >> 4) == mt) {
            // computed per m-tile from opaque copies of the ray: computed once
            // for all four, the 80 features would stay live in registers
            float oo[3] = {o[0], o[1], o[2]}, dd[3] = {d[0], d[1], d[2]};
            asm volatile("" : "+v"(oo[0]), "+v"(oo[1]), "+v"(oo[2]), "+v"(dd[0]), "+v"(dd[1]), "+v"(dd[2]));
            mlp_features(oo, dd, in + (l & 15) * MLP_K1);
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        mlp_layer<5, 4, false>(in, MLP_K1, w1, M.b, hid, nullptr);
        mlp_layer<4, 4, false>(hid, MLP_H, w2, M.b + MLP_H, hid, nullptr);
        mlp_layer<4, 4, false>(hid, MLP_H, w3, M.b + 2 * MLP_H, hid, nullptr);
        mlp_layer<4, 1, true>(hid, MLP_H, w4, M.b + 3 * MLP_H, nullptr, outf);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        if ((l >> 4) == mt) {
            t_out = outf[(l & 15) * 2];
            logit_out = outf[(l & 15) * 2 + 1];
        }
        __builtin_amdgcn_wave_barrier();
    }
}

}  // namespace mb3d
