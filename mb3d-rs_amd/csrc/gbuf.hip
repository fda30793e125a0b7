// gbuf.hip -- the G-buffer's file records: the card's 20-byte records (5
// u32, src/gpu.rs:3452-3463) to and from the 18-byte little-endian records
// of gbuffer::pack / unpack (src/gbuffer.rs:60-98), the pixel block of an
// .m3i (write_m3i, src/m3p.rs:1846-1853).
//
// Byte shuffling, HBM-bound: a block of 256 lanes takes 256 records per
// turn; both sides move through LDS as unit-stride dwords (256 x 18 bytes
// is 1152 dwords), so every global access is coalesced.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>

#include "../../include/mb3d_hip.h"
#include "host_common.h"
#include "paint_common.h"

namespace mb3d {

namespace {

constexpr int kRecBlock = 256;
constexpr int kRecWords = kRecBlock * 18 / 4;  // one block's file records in dwords

__device__ __forceinline__ void put16(unsigned char *b, uint32_t v) {
    b[0] = (unsigned char)(v & 0xFFu);
    b[1] = (unsigned char)((v >> 8) & 0xFFu);
}

__device__ __forceinline__ uint32_t get16(const unsigned char *b) { return (uint32_t)b[0] | ((uint32_t)b[1] << 8); }

__global__ void __launch_bounds__(kRecBlock) k_gbuf_pack(const uint32_t *__restrict__ gbuf, long long n,
                                                          unsigned char *__restrict__ out) {
    __shared__ uint32_t s_in[kRecBlock * 5];
    __shared__ uint32_t s_out[kRecWords];
    const int tid = threadIdx.x;
    const bool aligned = (reinterpret_cast<uintptr_t>(out) & 3u) == 0;
    for (long long base = (long long)blockIdx.x * kRecBlock; base < n; base += (long long)gridDim.x * kRecBlock) {
        const int m = (int)min((long long)kRecBlock, n - base);
        for (int i = tid; i < 5 * m; i += kRecBlock) s_in[i] = gbuf[5 * base + i];
        __syncthreads();
        unsigned char *sb = reinterpret_cast<unsigned char *>(s_out);
        if (tid < m) {
            const uint32_t *r = s_in + 5 * tid;
            unsigned char *b = sb + 18 * tid;
            put16(b + 0, r[0] & 0xFFFFu);   // normal[0]
            put16(b + 2, r[0] >> 16);       // normal[1]
            put16(b + 4, r[1] & 0xFFFFu);   // normal[2]
            put16(b + 6, r[2] & 0xFFFFu);   // zpos_fine, low half
            put16(b + 8, r[2] >> 16);       //            high half
            put16(b + 10, r[1] >> 16);      // shadow
            put16(b + 12, r[4] & 0xFFFFu);  // amb_shadow
            put16(b + 14, r[3] & 0xFFFFu);  // si_gradient
            put16(b + 16, r[3] >> 16);      // otrap
        }
        __syncthreads();
        unsigned char *dst = out + 18 * base;
        const int bytes = 18 * m;
        if (aligned) {
            uint32_t *d4 = reinterpret_cast<uint32_t *>(dst);
            for (int i = tid; i < bytes / 4; i += kRecBlock) d4[i] = s_out[i];
            for (int i = (bytes / 4) * 4 + tid; i < bytes; i += kRecBlock) dst[i] = sb[i];
        } else {
            for (int i = tid; i < bytes; i += kRecBlock) dst[i] = sb[i];
        }
        __syncthreads();
    }
}

__global__ void __launch_bounds__(kRecBlock) k_gbuf_unpack(const unsigned char *__restrict__ in, long long n,
                                                            uint32_t *__restrict__ gbuf) {
    __shared__ uint32_t s_in[kRecWords];
    __shared__ uint32_t s_out[kRecBlock * 5];
    const int tid = threadIdx.x;
    const bool aligned = (reinterpret_cast<uintptr_t>(in) & 3u) == 0;
    for (long long base = (long long)blockIdx.x * kRecBlock; base < n; base += (long long)gridDim.x * kRecBlock) {
        const int m = (int)min((long long)kRecBlock, n - base);
        const unsigned char *src = in + 18 * base;
        const int bytes = 18 * m;
        unsigned char *sb = reinterpret_cast<unsigned char *>(s_in);
        if (aligned) {
            const uint32_t *s4 = reinterpret_cast<const uint32_t *>(src);
            for (int i = tid; i < bytes / 4; i += kRecBlock) s_in[i] = s4[i];
            for (int i = (bytes / 4) * 4 + tid; i < bytes; i += kRecBlock) sb[i] = src[i];
        } else {
            for (int i = tid; i < bytes; i += kRecBlock) sb[i] = src[i];
        }
        __syncthreads();
        if (tid < m) {
            const unsigned char *b = sb + 18 * tid;
            uint32_t *r = s_out + 5 * tid;
            r[0] = get16(b + 0) | (get16(b + 2) << 16);
            r[1] = get16(b + 4) | (get16(b + 10) << 16);
            r[2] = get16(b + 6) | (get16(b + 8) << 16);
            r[3] = get16(b + 14) | (get16(b + 16) << 16);
            r[4] = get16(b + 12);
        }
        __syncthreads();
        for (int i = tid; i < 5 * m; i += kRecBlock) gbuf[5 * base + i] = s_out[i];
        __syncthreads();
    }
}

#define GBUF_TRY(expr)                                                                   \
    do {                                                                                 \
        hipError_t e_ = (expr);                                                          \
        if (e_ != hipSuccess) return set_error(MB3D_HIP_ERR_DEVICE, hipGetErrorString(e_)); \
    } while (0)

int convert_locked(bool pack, const void *in, long long n, void *out, int on_device) {
    if (lib_device() < 0) return set_error(MB3D_HIP_ERR_NOT_INIT, "mb3d_hip_init was not called");
    if (!in || !out || n < 0) return set_error(MB3D_HIP_ERR_INVALID, "bad argument");
    if (n == 0) return MB3D_HIP_OK;
    hipStream_t st = lib_stream();
    const size_t card = (size_t)n * 20, file = (size_t)n * 18;
    const size_t in_b = pack ? card : file, out_b = pack ? file : card;
    const void *d_in = in;
    void *d_out = out;
    if (!on_device) {
        int status = 0;
        char *work = static_cast<char *>(lib_work(in_b + out_b + 256, &status));
        if (!work) return status;
        d_in = work;
        d_out = work + ((in_b + 255) & ~(size_t)255);
        GBUF_TRY(hipMemcpyAsync(const_cast<void *>(d_in), in, in_b, hipMemcpyHostToDevice, st));
    }
    const long long blocks = (n + kRecBlock - 1) / kRecBlock;
    const int grid = (int)std::min<long long>(blocks, (long long)std::max(lib_cus(), 1) * 8);
    if (pack)
        hipLaunchKernelGGL(k_gbuf_pack, dim3(grid), dim3(kRecBlock), 0, st, static_cast<const uint32_t *>(d_in), n,
                           static_cast<unsigned char *>(d_out));
    else
        hipLaunchKernelGGL(k_gbuf_unpack, dim3(grid), dim3(kRecBlock), 0, st,
                           static_cast<const unsigned char *>(d_in), n, static_cast<uint32_t *>(d_out));
    GBUF_TRY(hipGetLastError());
    if (!on_device) GBUF_TRY(hipMemcpyAsync(out, d_out, out_b, hipMemcpyDeviceToHost, st));
    GBUF_TRY(hipStreamSynchronize(st));
    return MB3D_HIP_OK;
}

}  // namespace

}  // namespace mb3d

using namespace mb3d;

extern "C" int32_t mb3d_hip_gbuf_pack(const uint32_t *gbuf, int64_t n, uint8_t *records, int32_t on_device) {
    try {
        std::lock_guard<std::mutex> lk(lib_mutex());
        return convert_locked(true, gbuf, (long long)n, records, on_device);
    } catch (...) {
        return set_error(MB3D_HIP_ERR_DEVICE, "unexpected failure in mb3d_hip_gbuf_pack");
    }
}

extern "C" int32_t mb3d_hip_gbuf_unpack(const uint8_t *records, int64_t n, uint32_t *gbuf, int32_t on_device) {
    try {
        std::lock_guard<std::mutex> lk(lib_mutex());
        return convert_locked(false, records, (long long)n, gbuf, on_device);
    } catch (...) {
        return set_error(MB3D_HIP_ERR_DEVICE, "unexpected failure in mb3d_hip_gbuf_unpack");
    }
}
