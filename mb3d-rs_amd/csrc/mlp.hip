// mlp.hip -- the warm-start network: FP32 training kernels (forward, loss,
// backward, Adam), the BF16 MFMA inference kernel, and their C ABI
// (include/mb3d_hip.h, mb3d_hip_mlp_*).  No counterpart in the reference.
//
// Training runs the batch through one kernel per layer and pass: the
// layers are small (80 x 64 .. 64 x 2), so each kernel gives one output
// element to a thread; the weight gradients reduce over chunks of the batch,
// then sum the chunks in order, so that a gradient is the same on every call.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <random>
#include <vector>

#include "../../include/mb3d_hip.h"
#include "host_common.h"
#include "mlp_device.h"

static_assert(mb3d::MLP_PARAMS == MB3D_HIP_MLP_PARAMS, "parameter count");

struct mb3d_hip_mlp {
    float *params = nullptr;  // f32 master weights (MLP_PARAMS)
    float *grad = nullptr;
    float *m = nullptr, *v = nullptr;  // Adam moments
    uint16_t *packed = nullptr;        // bf16 MFMA tiles
    float *bias = nullptr;             // MLP_BIAS f32
    int step = 0;
};

namespace mb3d {

// offsets of the parameter blocks (W1, b1, W2, b2, W3, b3, W4, b4)
constexpr int OFF_W1 = 0, OFF_B1 = OFF_W1 + MLP_K1 * MLP_H, OFF_W2 = OFF_B1 + MLP_H, OFF_B2 = OFF_W2 + MLP_H * MLP_H,
              OFF_W3 = OFF_B2 + MLP_H, OFF_B3 = OFF_W3 + MLP_H * MLP_H, OFF_W4 = OFF_B3 + MLP_H,
              OFF_B4 = OFF_W4 + MLP_H * MLP_OUT;
constexpr float OVERSHOOT_WEIGHT = 4.0f;

// f32 features (the same recurrence as the inference's mlp_features)
__device__ void features_f32(const float *ray, float *f) {
    const float x[6] = {ray[0] * 0.25f, ray[1] * 0.25f, ray[2] * 0.25f, ray[3], ray[4], ray[5]};
    int c = 0;
    for (int i = 0; i < MLP_IN; i++) {
        float s, co;
        sincosf(3.14159265358979f * x[i], &s, &co);
        for (int k = 0; k < MLP_FREQ; k++) {
            f[c] = s;
            f[c + 1] = co;
            c += 2;
            const float s2 = 2.0f * s * co, c2 = 1.0f - 2.0f * s * s;
            s = s2;
            co = c2;
        }
    }
    for (int i = 0; i < MLP_IN; i++) f[c + i] = x[i];
    f[78] = 0.0f;
    f[79] = 0.0f;
}

__global__ void k_mlp_encode(const float *rays, long long n, float *X) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) features_f32(rays + 6 * i, X + MLP_K1 * i);
}

// Z = act(A) W + b for rows [0, n): one thread per output element; act is
// ReLU when the input is a previous layer's pre-activation
__global__ void k_mlp_fwd(const float *A, int relu_in, long long n, int K, int N, const float *W, const float *b,
                          float *Z) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * N) return;
    const long long r = t / N;
    const int c = (int)(t - r * N);
    const float *a = A + r * K;
    float acc = b[c];
    for (int k = 0; k < K; k++) {
        const float x = relu_in ? fmaxf(a[k], 0.0f) : a[k];
        acc = fmaf(x, W[k * N + c], acc);
    }
    Z[t] = acc;
}

// the loss of the outputs (mean over the n rays): BCE of the hit logit plus
// the asymmetric square of the distance error on hits; its gradient
__global__ void k_mlp_loss(const float *out, const float *tg, long long n, float *dout, float *loss_sum) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    float l = 0.0f;
    if (i < n) {
        const float t = tg[2 * i], hit = tg[2 * i + 1];
        const float th = out[2 * i], lg = out[2 * i + 1];
        const float p = 1.0f / (1.0f + expf(-lg));
        const float bce = fmaxf(lg, 0.0f) - lg * hit + log1pf(expf(-fabsf(lg)));
        const float e = th - t;
        const float a = e > 0.0f ? OVERSHOOT_WEIGHT : 1.0f;
        l = bce + hit * a * e * e;
        const float inv = 1.0f / (float)n;
        dout[2 * i] = hit * a * 2.0f * e * inv;
        dout[2 * i + 1] = (p - hit) * inv;
    }
    // a wave's sum, one atomic
    for (int o = 32; o > 0; o >>= 1) l += __shfl_xor(l, o, 64);
    if (__lane_id() == 0 && loss_sum) atomicAdd(loss_sum, l * (1.0f / (float)n));
}

// partial weight gradients over a chunk of MLP_ROWS rows (blockIdx.y):
// part[chunk][k * N + c] = sum act(A[r][k]) dZ[r][c], and for k = K the bias
// gradient sum dZ[r][c]; one thread per (k, c), c fastest, so the A reads are
// broadcasts and the dZ reads coalesced
constexpr int MLP_ROWS = 256;
__global__ void k_mlp_bwd_w(const float *A, int relu_in, const float *dZ, long long n, int K, int N, float *part) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int W = (K + 1) * N;
    if (t >= W) return;
    const int k = t / N, c = t - (t / N) * N;
    const long long r0 = (long long)blockIdx.y * MLP_ROWS;
    const long long r1 = r0 + MLP_ROWS < n ? r0 + MLP_ROWS : n;
    float acc = 0.0f;
    if (k < K) {
        for (long long r = r0; r < r1; r++) {
            const float a = A[r * K + k];
            acc = fmaf(relu_in ? fmaxf(a, 0.0f) : a, dZ[r * N + c], acc);
        }
    } else {
        for (long long r = r0; r < r1; r++) acc += dZ[r * N + c];
    }
    part[(size_t)blockIdx.y * W + t] = acc;
}

// the chunks' sum, in chunk order (deterministic), into gW (K x N) and gb
__global__ void k_mlp_bwd_sum(const float *part, int chunks, int K, int N, float *gW, float *gb) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int W = (K + 1) * N;
    if (t >= W) return;
    float acc = 0.0f;
    for (int q = 0; q < chunks; q++) acc += part[(size_t)q * W + t];
    if (t < K * N)
        gW[t] = acc;
    else
        gb[t - K * N] = acc;
}

// dA = (dZ W^T) * relu'(Zprev): one thread per (row, k)
__global__ void k_mlp_bwd_x(const float *dZ, const float *W, const float *Zprev, long long n, int K, int N,
                            float *dA) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * K) return;
    const long long r = t / K;
    const int k = (int)(t - r * K);
    float acc = 0.0f;
    for (int c = 0; c < N; c++) acc = fmaf(dZ[r * N + c], W[k * N + c], acc);
    dA[t] = Zprev[t] > 0.0f ? acc : 0.0f;
}

__global__ void k_adam(float *p, const float *g, float *m, float *v, int n, float lr, float b1, float b2, float eps,
                       float bc1, float bc2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
}

// keeps W1's two padding rows at zero
__global__ void k_mlp_clear_pad(float *p) {
    const int c = threadIdx.x;
    if (c < MLP_H) {
        p[OFF_W1 + 78 * MLP_H + c] = 0.0f;
        p[OFF_W1 + 79 * MLP_H + c] = 0.0f;
    }
}

// the MFMA path of the march, on its own: one wave per 64 rays
__global__ __launch_bounds__(256) void k_mlp_infer_mfma(MlpDev M, const float *rays, long long n, float *out) {
    __shared__ __attribute__((aligned(16))) char lds[4 * MLP_LDS_BYTES];
    __shared__ __attribute__((aligned(16))) mlp_s4 wl[MLP_TILES * 64];
    mlp_stage_weights(M, wl);
    __syncthreads();
    const int wave = threadIdx.x >> 6;
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long j = i < n ? i : n - 1;
    const float o[3] = {rays[6 * j], rays[6 * j + 1], rays[6 * j + 2]};
    const float d[3] = {rays[6 * j + 3], rays[6 * j + 4], rays[6 * j + 5]};
    float t, lg;
    mlp_wave_infer(M, wl, o, d, lds + wave * MLP_LDS_BYTES, t, lg);
    if (i < n) {
        out[2 * i] = t;
        out[2 * i + 1] = lg;
    }
}

// host: bf16 rounding and the MFMA tile packing of the weights
uint16_t host_bf16(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

// The MFMA packing of W (K x N, row-major) as the A operand of
// out^T = W^T in^T (mlp_device.h): per tile (kt, nt), lane l = 16 q + m,
// `per` values of W[k][16 nt + m], the inputs k of the lane's K slots
// (`slot(kt, q, i)`), zero beyond K or N.
template <class Slot>
void pack_tiles(const float *W, int K, int N, int KT, int NT, int per, Slot slot, uint16_t *dst) {
    for (int kt = 0; kt < KT; kt++)
        for (int nt = 0; nt < NT; nt++)
            for (int l = 0; l < 64; l++)
                for (int i = 0; i < per; i++) {
                    const int k = slot(kt, l / 16, i), c = 16 * nt + l % 16;
                    const float w = (k < K && c < N) ? W[k * N + c] : 0.0f;
                    dst[((kt * NT + nt) * 64 + l) * per + i] = host_bf16(w);
                }
}

// the first layer's K = 32 slots: features 32 kt + 8 q + i (read from LDS
// in that order); its K = 16 tile: features 64 + 4 q + i
int slot_l1a(int kt, int q, int i) { return 32 * kt + 8 * q + i; }
int slot_l1b(int, int q, int i) { return 64 + 4 * q + i; }
// the K = 32 slots of a layer fed by the previous layer's D tiles 2 kt
// (elements 0..3) and 2 kt + 1 (elements 4..7)
int slot_chain(int kt, int q, int i) { return 32 * kt + (i < 4 ? 4 * q + i : 16 + 4 * q + (i - 4)); }

int repack(mb3d_hip_mlp *net, hipStream_t s) {
    std::vector<float> p(MLP_PARAMS);
    if (hipMemcpyAsync(p.data(), net->params, MLP_PARAMS * 4, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        return set_error(MB3D_HIP_ERR_DEVICE, "copying the network's weights failed");
    std::vector<uint16_t> w(MLP_PACKED_U16);
    uint16_t *q = w.data();
    pack_tiles(p.data() + OFF_W1, MLP_K1, MLP_H, 2, 4, 8, slot_l1a, q + MLP_OFF_L1A);
    pack_tiles(p.data() + OFF_W1, MLP_K1, MLP_H, 1, 4, 4, slot_l1b, q + MLP_OFF_L1B);
    pack_tiles(p.data() + OFF_W2, MLP_H, MLP_H, 2, 4, 8, slot_chain, q + MLP_OFF_L2);
    pack_tiles(p.data() + OFF_W3, MLP_H, MLP_H, 2, 4, 8, slot_chain, q + MLP_OFF_L3);
    pack_tiles(p.data() + OFF_W4, MLP_H, MLP_OUT, 2, 1, 8, slot_chain, q + MLP_OFF_L4);
    std::vector<float> b(MLP_BIAS, 0.0f);
    std::memcpy(b.data(), p.data() + OFF_B1, MLP_H * 4);
    std::memcpy(b.data() + MLP_H, p.data() + OFF_B2, MLP_H * 4);
    std::memcpy(b.data() + 2 * MLP_H, p.data() + OFF_B3, MLP_H * 4);
    std::memcpy(b.data() + 3 * MLP_H, p.data() + OFF_B4, MLP_OUT * 4);
    if (hipMemcpyAsync(net->packed, w.data(), w.size() * 2, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(net->bias, b.data(), b.size() * 4, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        return set_error(MB3D_HIP_ERR_DEVICE, "uploading the packed weights failed");
    return 0;
}

// device scratch of one training or gradient call
struct TrainBufs {
    float *X = nullptr, *tg = nullptr, *Z1 = nullptr, *Z2 = nullptr, *Z3 = nullptr, *O = nullptr;
    float *dO = nullptr, *d3 = nullptr, *d2 = nullptr, *d1 = nullptr, *loss = nullptr, *part = nullptr;
    std::vector<void *> owned;
    ~TrainBufs() {
        for (void *p : owned) (void)hipFree(p);
    }
    int alloc(float **p, size_t n) {
        if (hipMalloc((void **)p, std::max<size_t>(n, 1) * 4) != hipSuccess)
            return set_error(MB3D_HIP_ERR_DEVICE, "hipMalloc failed for the network's training");
        owned.push_back(*p);
        return 0;
    }
};

inline int blocks(long long n) { return (int)((n + 255) / 256); }

// the weight and bias gradients of one layer
void bwd_w(TrainBufs &T, const float *A, int relu_in, const float *dZ, long long b, int K, int N, float *gW,
           float *gb, hipStream_t s) {
    const int chunks = (int)((b + MLP_ROWS - 1) / MLP_ROWS);
    hipLaunchKernelGGL(k_mlp_bwd_w, dim3(blocks((K + 1) * N), chunks), dim3(256), 0, s, A, relu_in, dZ, b, K, N,
                       T.part);
    hipLaunchKernelGGL(k_mlp_bwd_sum, dim3(blocks((K + 1) * N)), dim3(256), 0, s, T.part, chunks, K, N, gW, gb);
}

// forward + loss + backward of rows [r0, r0 + b) of the encoded set into net->grad
int fwd_bwd(mb3d_hip_mlp *net, TrainBufs &T, long long r0, long long b, float *loss_slot, hipStream_t s) {
    const float *P = net->params;
    float *G = net->grad;
    const float *X = T.X + r0 * MLP_K1;
    const float *tg = T.tg + r0 * 2;
    hipLaunchKernelGGL(k_mlp_fwd, dim3(blocks(b * MLP_H)), dim3(256), 0, s, X, 0, b, MLP_K1, MLP_H, P + OFF_W1,
                       P + OFF_B1, T.Z1);
    hipLaunchKernelGGL(k_mlp_fwd, dim3(blocks(b * MLP_H)), dim3(256), 0, s, T.Z1, 1, b, MLP_H, MLP_H, P + OFF_W2,
                       P + OFF_B2, T.Z2);
    hipLaunchKernelGGL(k_mlp_fwd, dim3(blocks(b * MLP_H)), dim3(256), 0, s, T.Z2, 1, b, MLP_H, MLP_H, P + OFF_W3,
                       P + OFF_B3, T.Z3);
    hipLaunchKernelGGL(k_mlp_fwd, dim3(blocks(b * MLP_OUT)), dim3(256), 0, s, T.Z3, 1, b, MLP_H, MLP_OUT,
                       P + OFF_W4, P + OFF_B4, T.O);
    hipLaunchKernelGGL(k_mlp_loss, dim3(blocks(b)), dim3(256), 0, s, T.O, tg, b, T.dO, loss_slot);
    bwd_w(T, T.Z3, 1, T.dO, b, MLP_H, MLP_OUT, G + OFF_W4, G + OFF_B4, s);
    hipLaunchKernelGGL(k_mlp_bwd_x, dim3(blocks(b * MLP_H)), dim3(256), 0, s, T.dO, P + OFF_W4, T.Z3, b, MLP_H,
                       MLP_OUT, T.d3);
    bwd_w(T, T.Z2, 1, T.d3, b, MLP_H, MLP_H, G + OFF_W3, G + OFF_B3, s);
    hipLaunchKernelGGL(k_mlp_bwd_x, dim3(blocks(b * MLP_H)), dim3(256), 0, s, T.d3, P + OFF_W3, T.Z2, b, MLP_H,
                       MLP_H, T.d2);
    bwd_w(T, T.Z1, 1, T.d2, b, MLP_H, MLP_H, G + OFF_W2, G + OFF_B2, s);
    hipLaunchKernelGGL(k_mlp_bwd_x, dim3(blocks(b * MLP_H)), dim3(256), 0, s, T.d2, P + OFF_W2, T.Z1, b, MLP_H,
                       MLP_H, T.d1);
    bwd_w(T, X, 0, T.d1, b, MLP_K1, MLP_H, G + OFF_W1, G + OFF_B1, s);
    if (hipGetLastError() != hipSuccess) return set_error(MB3D_HIP_ERR_DEVICE, "a training kernel failed to launch");
    return 0;
}

int setup_set(TrainBufs &T, const float *rays, const float *targets, long long n, long long b, hipStream_t s) {
    float *d_rays = nullptr;
    if (int e = T.alloc(&d_rays, (size_t)n * 6)) return e;
    if (int e = T.alloc(&T.X, (size_t)n * MLP_K1)) return e;
    if (int e = T.alloc(&T.tg, (size_t)n * 2)) return e;
    if (int e = T.alloc(&T.Z1, (size_t)b * MLP_H)) return e;
    if (int e = T.alloc(&T.Z2, (size_t)b * MLP_H)) return e;
    if (int e = T.alloc(&T.Z3, (size_t)b * MLP_H)) return e;
    if (int e = T.alloc(&T.O, (size_t)b * 2)) return e;
    if (int e = T.alloc(&T.dO, (size_t)b * 2)) return e;
    if (int e = T.alloc(&T.d3, (size_t)b * MLP_H)) return e;
    if (int e = T.alloc(&T.d2, (size_t)b * MLP_H)) return e;
    if (int e = T.alloc(&T.d1, (size_t)b * MLP_H)) return e;
    if (int e = T.alloc(&T.part, (size_t)((b + MLP_ROWS - 1) / MLP_ROWS) * (MLP_K1 + 1) * MLP_H)) return e;
    if (hipMemcpyAsync(d_rays, rays, (size_t)n * 24, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(T.tg, targets, (size_t)n * 8, hipMemcpyHostToDevice, s) != hipSuccess)
        return set_error(MB3D_HIP_ERR_DEVICE, "uploading the training set failed");
    hipLaunchKernelGGL(k_mlp_encode, dim3(blocks(n)), dim3(256), 0, s, d_rays, n, T.X);
    return 0;
}

}  // namespace mb3d

using namespace mb3d;

extern "C" int32_t mb3d_hip_mlp_create(uint64_t seed, mb3d_hip_mlp **net) {
    if (!net) return set_error(MB3D_HIP_ERR_INVALID, "net is NULL");
    *net = nullptr;
    std::lock_guard<std::mutex> lk(lib_mutex());
    hipStream_t s = lib_stream();
    if (!s) return set_error(MB3D_HIP_ERR_NOT_INIT, "mb3d_hip_init was not called");
    mb3d_hip_mlp *h = new mb3d_hip_mlp();
    auto fail = [&](const char *m) {
        mb3d_hip_mlp_destroy(h);
        return set_error(MB3D_HIP_ERR_DEVICE, m);
    };
    if (hipMalloc((void **)&h->params, MLP_PARAMS * 4) != hipSuccess ||
        hipMalloc((void **)&h->grad, MLP_PARAMS * 4) != hipSuccess ||
        hipMalloc((void **)&h->m, MLP_PARAMS * 4) != hipSuccess ||
        hipMalloc((void **)&h->v, MLP_PARAMS * 4) != hipSuccess ||
        hipMalloc((void **)&h->packed, MLP_PACKED_U16 * 2) != hipSuccess ||
        hipMalloc((void **)&h->bias, MLP_BIAS * 4) != hipSuccess)
        return fail("hipMalloc failed for the network");
    // He initialisation of the weights, zero biases
    std::vector<float> p(MLP_PARAMS, 0.0f);
    std::mt19937_64 rng(seed);
    auto fill = [&](int off, int K, int N) {
        std::normal_distribution<float> nd(0.0f, std::sqrt(2.0f / (float)K));
        for (int i = 0; i < K * N; i++) p[off + i] = nd(rng);
    };
    fill(OFF_W1, MLP_K1, MLP_H);
    for (int k = 78; k < MLP_K1; k++)
        for (int c = 0; c < MLP_H; c++) p[OFF_W1 + k * MLP_H + c] = 0.0f;
    fill(OFF_W2, MLP_H, MLP_H);
    fill(OFF_W3, MLP_H, MLP_H);
    fill(OFF_W4, MLP_H, MLP_OUT);
    std::vector<float> zero(MLP_PARAMS, 0.0f);
    if (hipMemcpyAsync(h->params, p.data(), MLP_PARAMS * 4, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(h->m, zero.data(), MLP_PARAMS * 4, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(h->v, zero.data(), MLP_PARAMS * 4, hipMemcpyHostToDevice, s) != hipSuccess)
        return fail("initialising the network failed");
    if (int e = repack(h, s)) {
        mb3d_hip_mlp_destroy(h);
        return e;
    }
    *net = h;
    return MB3D_HIP_OK;
}

extern "C" void mb3d_hip_mlp_destroy(mb3d_hip_mlp *h) {
    if (!h) return;
    for (void *p : {(void *)h->params, (void *)h->grad, (void *)h->m, (void *)h->v, (void *)h->packed, (void *)h->bias})
        if (p) (void)hipFree(p);
    delete h;
}

extern "C" int32_t mb3d_hip_mlp_params(mb3d_hip_mlp *h, float *get, const float *set, int64_t n) {
    if (!h || n != MLP_PARAMS) return set_error(MB3D_HIP_ERR_INVALID, "bad network or parameter count");
    std::lock_guard<std::mutex> lk(lib_mutex());
    hipStream_t s = lib_stream();
    if (!s) return set_error(MB3D_HIP_ERR_NOT_INIT, "mb3d_hip_init was not called");
    if (get) {
        if (hipMemcpyAsync(get, h->params, MLP_PARAMS * 4, hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess)
            return set_error(MB3D_HIP_ERR_DEVICE, "reading the weights failed");
    }
    if (set) {
        if (hipMemcpyAsync(h->params, set, MLP_PARAMS * 4, hipMemcpyHostToDevice, s) != hipSuccess)
            return set_error(MB3D_HIP_ERR_DEVICE, "writing the weights failed");
        hipLaunchKernelGGL(k_mlp_clear_pad, dim3(1), dim3(64), 0, s, h->params);
        return repack(h, s);
    }
    return MB3D_HIP_OK;
}

extern "C" int32_t mb3d_hip_mlp_grad(mb3d_hip_mlp *h, const float *rays, const float *targets, int64_t n,
                                     float *loss, float *grad) {
    if (!h || !rays || !targets || n <= 0) return set_error(MB3D_HIP_ERR_INVALID, "bad argument");
    std::lock_guard<std::mutex> lk(lib_mutex());
    hipStream_t s = lib_stream();
    if (!s) return set_error(MB3D_HIP_ERR_NOT_INIT, "mb3d_hip_init was not called");
    TrainBufs T;
    if (int e = setup_set(T, rays, targets, n, n, s)) return e;
    if (int e = T.alloc(&T.loss, 1)) return e;
    if (hipMemsetAsync(T.loss, 0, 4, s) != hipSuccess) return set_error(MB3D_HIP_ERR_DEVICE, "memset failed");
    if (int e = fwd_bwd(h, T, 0, n, T.loss, s)) return e;
    if (loss && hipMemcpyAsync(loss, T.loss, 4, hipMemcpyDeviceToHost, s) != hipSuccess)
        return set_error(MB3D_HIP_ERR_DEVICE, "copy failed");
    if (grad && hipMemcpyAsync(grad, h->grad, MLP_PARAMS * 4, hipMemcpyDeviceToHost, s) != hipSuccess)
        return set_error(MB3D_HIP_ERR_DEVICE, "copy failed");
    if (hipStreamSynchronize(s) != hipSuccess) return set_error(MB3D_HIP_ERR_DEVICE, "the gradient failed");
    return MB3D_HIP_OK;
}

extern "C" int32_t mb3d_hip_mlp_train(mb3d_hip_mlp *h, const float *rays, const float *targets, int64_t n,
                                      int32_t steps, int32_t batch, float lr, float *loss_curve) {
    if (!h || !rays || !targets || n <= 0 || steps <= 0 || batch <= 0)
        return set_error(MB3D_HIP_ERR_INVALID, "bad argument");
    std::lock_guard<std::mutex> lk(lib_mutex());
    hipStream_t s = lib_stream();
    if (!s) return set_error(MB3D_HIP_ERR_NOT_INIT, "mb3d_hip_init was not called");
    const long long b = std::min<long long>(batch, n);
    TrainBufs T;
    if (int e = setup_set(T, rays, targets, n, b, s)) return e;
    if (int e = T.alloc(&T.loss, (size_t)steps)) return e;
    if (hipMemsetAsync(T.loss, 0, (size_t)steps * 4, s) != hipSuccess)
        return set_error(MB3D_HIP_ERR_DEVICE, "memset failed");
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    long long r0 = 0;
    for (int st = 0; st < steps; st++) {
        if (r0 + b > n) r0 = 0;
        if (int e = fwd_bwd(h, T, r0, b, T.loss + st, s)) return e;
        r0 += b;
        h->step++;
        const float frac = (float)st / (float)steps;
        const float lr_t = lr * (0.05f + 0.95f * 0.5f * (1.0f + std::cos(3.14159265f * frac)));
        const float bc1 = 1.0f - std::pow(b1, (float)h->step), bc2 = 1.0f - std::pow(b2, (float)h->step);
        hipLaunchKernelGGL(k_adam, dim3(blocks(MLP_PARAMS)), dim3(256), 0, s, h->params, h->grad, h->m, h->v,
                           (int)MLP_PARAMS, lr_t, b1, b2, eps, bc1, bc2);
        hipLaunchKernelGGL(k_mlp_clear_pad, dim3(1), dim3(64), 0, s, h->params);
    }
    if (hipGetLastError() != hipSuccess) return set_error(MB3D_HIP_ERR_DEVICE, "a training kernel failed");
    if (loss_curve && hipMemcpyAsync(loss_curve, T.loss, (size_t)steps * 4, hipMemcpyDeviceToHost, s) != hipSuccess)
        return set_error(MB3D_HIP_ERR_DEVICE, "copy failed");
    if (hipStreamSynchronize(s) != hipSuccess) return set_error(MB3D_HIP_ERR_DEVICE, "training failed");
    return repack(h, s);
}

extern "C" int32_t mb3d_hip_mlp_infer(mb3d_hip_mlp *h, const float *rays, int64_t n, int32_t bf16_mfma, float *out) {
    if (!h || !rays || !out || n <= 0) return set_error(MB3D_HIP_ERR_INVALID, "bad argument");
    std::lock_guard<std::mutex> lk(lib_mutex());
    hipStream_t s = lib_stream();
    if (!s) return set_error(MB3D_HIP_ERR_NOT_INIT, "mb3d_hip_init was not called");
    TrainBufs T;
    float *d_rays = nullptr, *d_out = nullptr;
    if (int e = T.alloc(&d_rays, (size_t)n * 6)) return e;
    if (int e = T.alloc(&d_out, (size_t)n * 2)) return e;
    if (hipMemcpyAsync(d_rays, rays, (size_t)n * 24, hipMemcpyHostToDevice, s) != hipSuccess)
        return set_error(MB3D_HIP_ERR_DEVICE, "copy failed");
    if (bf16_mfma) {
        MlpDev M;
        std::memset(&M, 0, sizeof M);
        M.w = h->packed;
        M.b = h->bias;
        M.enabled = 1;
        hipLaunchKernelGGL(k_mlp_infer_mfma, dim3(blocks(n)), dim3(256), 0, s, M, d_rays, (long long)n, d_out);
    } else {
        if (int e = T.alloc(&T.X, (size_t)n * MLP_K1)) return e;
        if (int e = T.alloc(&T.Z1, (size_t)n * MLP_H)) return e;
        if (int e = T.alloc(&T.Z2, (size_t)n * MLP_H)) return e;
        if (int e = T.alloc(&T.Z3, (size_t)n * MLP_H)) return e;
        const float *P = h->params;
        hipLaunchKernelGGL(k_mlp_encode, dim3(blocks(n)), dim3(256), 0, s, d_rays, (long long)n, T.X);
        hipLaunchKernelGGL(k_mlp_fwd, dim3(blocks(n * MLP_H)), dim3(256), 0, s, T.X, 0, (long long)n, MLP_K1, MLP_H,
                           P + OFF_W1, P + OFF_B1, T.Z1);
        hipLaunchKernelGGL(k_mlp_fwd, dim3(blocks(n * MLP_H)), dim3(256), 0, s, T.Z1, 1, (long long)n, MLP_H, MLP_H,
                           P + OFF_W2, P + OFF_B2, T.Z2);
        hipLaunchKernelGGL(k_mlp_fwd, dim3(blocks(n * MLP_H)), dim3(256), 0, s, T.Z2, 1, (long long)n, MLP_H, MLP_H,
                           P + OFF_W3, P + OFF_B3, T.Z3);
        hipLaunchKernelGGL(k_mlp_fwd, dim3(blocks(n * MLP_OUT)), dim3(256), 0, s, T.Z3, 1, (long long)n, MLP_H,
                           MLP_OUT, P + OFF_W4, P + OFF_B4, d_out);
    }
    if (hipGetLastError() != hipSuccess) return set_error(MB3D_HIP_ERR_DEVICE, "inference failed to launch");
    if (hipMemcpyAsync(out, d_out, (size_t)n * 8, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        return set_error(MB3D_HIP_ERR_DEVICE, "inference failed");
    return MB3D_HIP_OK;
}

// the device form of a network for the march
namespace mb3d {
void mlp_dev(const mb3d_hip_mlp *h, const mb3d_hip_options &o, MlpDev &M) {
    std::memset(&M, 0, sizeof M);
    if (!h) return;
    M.w = h->packed;
    M.b = h->bias;
    // a hit threshold of 0 selects 0.5, above 1 turns hit jumps off; miss
    // jumps (unverified) only with a positive threshold
    M.conf_hit = o.mlp_conf_hit != 0.0f ? o.mlp_conf_hit : 0.5f;
    M.conf_miss = o.mlp_conf_miss > 0.0f ? o.mlp_conf_miss : -1.0f;
    M.margin_rel = o.mlp_margin_rel > 0.0f ? o.mlp_margin_rel : 0.15f;
    M.margin_abs = o.mlp_margin_abs;
    M.enabled = 1;
    if (o.mlp_cone_gate > 0.0f && o.mlp_cone_gate < 1.0f && o.mlp_cone != 0) {
        M.gate = 1;
        M.gate_logit = std::log(o.mlp_cone_gate / (1.0f - o.mlp_cone_gate));
    }
}
}  // namespace mb3d
