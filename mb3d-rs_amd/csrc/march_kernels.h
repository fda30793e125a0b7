// march_kernels.h -- the device code of the DE ray marcher for gfx950: the
// kernels every pass runs, as templates over the formula kind and DE mode.
// march.hip instantiates the FP32 kinds' kernels and holds the C ABI;
// march_gen.hip instantiates the general kind's (K_GEN), so that the two
// compile in parallel.
//
// Passes of one call (all on one stream, no host round trip in between):
//  0. k_mlp_propose  the warm-start network's proposal per work item (opt-in)
//  1. k_march32  persistent, wave64 lanes refilled from a tile queue: the
//                FP32-iterate march (march_pixel_from + march_from,
//                src/calc.rs:2231-2322,2533-2657) up to the guard band
//                around the DE stop; background rays finish here, rays near
//                the surface are handed to pass 2 through a compacted list.
//  2. k_finish64 the reference's f64 marcher from the hand-off state: the
//                stop decision and bisection (src/calc.rs:1706-1792), up to
//                the refined surface point, listed for pass 3.
//  3. k_normals  per refined surface point: normals (1813-1858), colour and
//                depth (2471-2517, 2194-2205), the G-buffer words.
//  4. k_shadow   per surface pixel: hs_surface_start + the march towards each
//                light (src/calc.rs:2741-2899, without the row shortcut).
//  5. k_deao     per surface pixel: surface_point + DE AO rays
//                (src/calc.rs:2157-2192, src/deao.rs:268-418).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstddef>
#include <cstdint>

#include "../../include/mb3d_hip.h"
#include "device_de.h"
#include "mlp_device.h"

namespace mb3d {

// ---------------------------------------------------------------------------
// device buffers of one call

// State of a ray handed from the FP32 march to the FP64 finish.
struct Cand {
    double c[3], v[3];
    double mzz, last_step, last_de;
    float rsf, step_count;
    int32_t pix, flags, nsteps, seed;
};
enum : int32_t { CF_START = 1, CF_FIRST = 2, CF_STEPPED = 4 };

// A refined surface point, from the finish to the normals pass, written and
// read in list order (coalesced): the point, its depth and DE stop (the DE
// stop of its last update: bin_search_it moves the point without updating
// it), the step count, the pixel, and the iteration count with the flags.
struct SurfRec {
    double c[3];
    double mzz;
    float ms, step_count;
    int32_t pix;
    int32_t info;  // iteration count | SR_DELIM | frame << 20
};
enum : int32_t { SR_ITERS = 0xFFFF, SR_DELIM = 1 << 16 };  // SR_DELIM: the stop was the DE's, not the limit's

// counter words of Bufs::ctr
// Each queue counter has a 256-byte line of its own: device-wide atomics on
// one line serialise, and two counters of one kernel sharing a line would
// queue behind each other.
enum : int {
    C_LINE = 32,             // words per counter line
    C_QUEUE = 0 * C_LINE,    // next work item of the march queue
    C_CAND = 1 * C_LINE,     // hand-off list length
    C_HITS = 2 * C_LINE,     // surface pixel list length
    C_FINQ = 3 * C_LINE,     // next item of the f64 finish
    C_POSTQ = 4 * C_LINE,    // next item of the shadow pass
    C_AOQ = 5 * C_LINE,      // next item of the DE AO pass
    C_SURF = 6 * C_LINE,     // refined surface list length (the normals pass's items)
    C_CONEQ = 7 * C_LINE,    // next tile of the cone rays
    C_LIVE = 8 * C_LINE,     // tiles the cone rays leave to the march (k_cone_bg's list)
    C_STATS = 9 * C_LINE,    // 19 work counters from here (the 16th: the network's rays)
    C_WORDS = C_STATS + 19
};

struct Bufs {
    float *depth;
    double *depth64;
    uint16_t *iters;
    float *normals;
    uint32_t *gbuf;
    Cand *cand;
    int32_t *hits;
    SurfRec *surf;  // the refined surface points (pix -1: a chunk's empty tail)
    unsigned long long *ctr;
    // the warm start: the network's proposal per work item of the queue
    // (hit distance / 4, hit logit)
    float2 *prop;
    // the frames' ray tables (k_ray_tab), read by the FP32 march and the
    // warm start's proposals
    const double2 *rtab;
    // the cone rays' results, one per tile of the queue (k_cone)
    struct ConeRes *cone;
};

// A tile's cone ray (k_cone): the depth up to which every ray of the tile
// is free of the surface (+inf: the cone passed the end of the view) and the
// dynamic fog's step count up to there.  `live` is not the tile's: record s
// holds entry s of the list of tiles the march takes (k_cone_bg writes it
// after reading the records' depths and counts; the tile's index in the
// queue | its frame << 26).
struct ConeRes {
    double t;
    float count;
    int32_t live;
};
constexpr int LIVE_TILE_BITS = 26;

// Fills the ray tables of the frames `tab[0..n)` (one frame: `tab` = &P):
// block row blockIdx.y takes frame blockIdx.y's width + rows entries.
static __global__ void k_ray_tab(const DevP *__restrict__ tab, DevP one, double2 *rtab) {
    const DevP &P = tab ? tab[blockIdx.y] : one;
    const int n = P.width + P.rows;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        double2 e;
        ray_tab_entry(P, i, e);
        rtab[P.rt_base + i] = e;
    }
}

// DE AO ray set (build_rays, src/deao.rs:75-105), computed on the host
// One ray set of the DE AO in the frame of a surface whose normal is -z
struct AoSet {
    float dir[33][3];
    int count;
    int pad[3];
};

struct AoRays {
    float dir[33][3];
    int count;
    float d_step_mul, d_min_a_dif, correction_weight, max_len;
    // the sets per dither cell: sets[(y % side) * side + x % side] (side =
    // dither + 1; side 1: the set above, as sets[0])
    const AoSet *sets;
    int side;
    // the random first step: each object pixel's generator state before its
    // first ray (k_ao_seed), else NULL
    const uint32_t *rnd;
};

struct ShadowJob {
    int n;
    int bit[MB3D_HIP_MAX_LIGHTS];
    double vec[MB3D_HIP_MAX_LIGHTS][3];
    float max_len_mul;
};

// ---------------------------------------------------------------------------
// wave helpers

__device__ __forceinline__ int lane_id() { return __lane_id(); }

__device__ __forceinline__ unsigned long long wave_sum(unsigned long long v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ inline void flush_counters(const Cnt &c, unsigned long long *ctr) {
    const unsigned long long v[18] = {c.steps,  c.de32, c.de64,     c.p8_32, c.box_32,    c.p8_64, c.box_64,
                                      c.hits,   c.overflow, c.jumps, c.rollbacks, c.oth_64, c.nrm_32,
                                      c.cone_rays, c.cone_bg, c.cone_its, c.cone_des, c.steps64};
#pragma unroll
    for (int k = 0; k < 18; k++) {
        const unsigned long long s = wave_sum(v[k]);
        // (word 15 is the network's ray count, k_mlp_propose's own)
        if (lane_id() == 0 && s) atomicAdd(&ctr[C_STATS + (k < 15 ? k : k + 1)], s);
    }
}

// The value of the wave's first active lane, as a wave-uniform (scalar)
// value: broadcasts an atomic's result from the lane that made it (always
// the first active lane here), so that the compiler sees the loops whose exit
// depends on it as uniform.
__device__ __forceinline__ unsigned long long first_lane_u64(unsigned long long v) {
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    return ((unsigned long long)hi << 32) | lo;
}

// appends one entry per lane with `want` set: returns the lane's slot
__device__ __forceinline__ long long wave_append(bool want, unsigned long long *counter) {
    const unsigned long long m = __ballot(want);
    if (!m) return -1;
    // the first active lane makes the atomic (first_lane_u64 reads it)
    const int leader = __ffsll((long long)__ballot(true)) - 1;
    // every active lane makes the atomic, adding 0 but the first (the
    // compiler folds them into one): no lane-specific branch
    const unsigned long long base =
        first_lane_u64(atomicAdd(counter, lane_id() == leader ? (unsigned long long)__popcll(m) : 0ull));
    if (!want) return -1;
    return (long long)(base + __popcll(m & ((1ull << lane_id()) - 1ull)));
}

// A wave's share of a global counter, taken `chunk` items at a time, so that
// the lanes of a persistent wave get their next items without an atomic per
// turn (a device-wide atomic on one address serialises across the XCDs).
struct WaveChunk {
    long long base;  // next item of the wave's chunk (the same in every lane)
    int left;        // items left in the chunk
    bool dry;        // the counter passed its end
};

__device__ __forceinline__ void chunk_init(WaveChunk &q) {
    q.base = 0;
    q.left = 0;
    q.dry = false;
}

__device__ __forceinline__ unsigned long long lanemask_lt() { return (1ull << lane_id()) - 1ull; }

struct NoHook {
    __device__ void operator()(long long, int) const {}
};

// Hands the lanes with `want` the next items of a queue of `total` items
// (`total` < 0: no end, an allocation).  Call with the whole wave active;
// `hook(base, count)` runs, with the whole wave active, for each new chunk
// the wave takes.  Returns the lane's item, or -1 (the queue is exhausted).
template <typename Hook = NoHook>
__device__ __forceinline__ long long wave_take(bool want, unsigned long long *ctr, WaveChunk &q, int chunk,
                                               long long total, Hook &&hook = Hook()) {
    long long item = -1;
    bool pending = want;
    const int leader = __ffsll((long long)__ballot(true)) - 1;
    for (;;) {
        const unsigned long long m = __ballot(pending);
        if (!m || q.dry) break;
        if (q.left == 0) {
            const unsigned long long b =
                first_lane_u64(atomicAdd(ctr, lane_id() == leader ? (unsigned long long)chunk : 0ull));
            if (total >= 0 && (long long)b >= total) {
                q.dry = true;
                break;
            }
            q.base = (long long)b;
            q.left = total >= 0 ? (int)min((long long)chunk, total - (long long)b) : chunk;
            hook(q.base, q.left);
        }
        const int rank = __popcll(m & lanemask_lt());
        const int take = min(__popcll(m), q.left);
        if (pending && rank < take) {
            item = q.base + rank;
            pending = false;
        }
        q.base += take;
        q.left -= take;
    }
    return item;
}

__device__ inline void write_background(const DevP &P, const Bufs &B, int pix, float step_count) {
    const uint32_t sh = (uint32_t)rtef(fminf(fmaxf(step_count, 0.0f), 1023.0f));
    if (B.depth) B.depth[pix] = __builtin_inff();
    if (B.depth64) B.depth64[pix] = __builtin_inf();
    if (B.iters) B.iters[pix] = 0;
    if (B.normals) {
        B.normals[3 * pix] = 0.0f;
        B.normals[3 * pix + 1] = 0.0f;
        B.normals[3 * pix + 2] = 0.0f;
    }
    if (B.gbuf) {
        uint32_t *g = B.gbuf + 5 * (size_t)pix;
        g[0] = 0;
        g[1] = sh << 16;
        g[2] = 32768u << 16;
        g[3] = 0;
        g[4] = 5000;
    }
}

// pixel_seed (src/calc.rs:1125-1140)
__device__ inline int32_t dev_pixel_seed(int64_t y, int64_t x) {
    uint32_t s = (uint32_t)(int32_t)(0x24563487LL + (y + 1) * 0x324594A1LL);
    uint32_t mul = 214013u, add = 2531011u, am = 1u, aa = 0u;
    for (uint64_t n = x > 0 ? (uint64_t)x : 0; n; n >>= 1) {
        if (n & 1) {
            aa = aa * mul + add;
            am *= mul;
        }
        add = add * mul + add;
        mul *= mul;
    }
    return (int32_t)(am * s + aa);
}

// ---------------------------------------------------------------------------
// pass 1: the FP32 march, persistent, lanes refilled one by one

// A batch's frame table: frame f's parameters at base + f * STRIDE bytes.
// In global memory the records are whole DevPs; the march and finish
// kernels stage their cores (the fields before DevP::opt2, all that the
// batched kinds read) in LDS, DEVP_CORE bytes apart.
constexpr int DEVP_CORE = (int)offsetof(DevP, opt2);
static_assert(DEVP_CORE % 8 == 0, "the core of DevP in 8-byte words");
template <int STRIDE>
struct Frames {
    const char *base;
    __device__ __forceinline__ const DevP &operator[](int f) const {
        return *reinterpret_cast<const DevP *>(base + (size_t)f * STRIDE);
    }
};
using LdsFrames = Frames<DEVP_CORE>;
using GlobalFrames = Frames<(int)sizeof(DevP)>;

// The cores of a batch's frames copied to LDS (the dynamic LDS of the
// launch, nframes x DEVP_CORE bytes): the march and finish read a frame's
// parameters in their inner loops, and from global memory every read is a
// vector load (the frame differs per lane).  All threads of the block call
// it.
__device__ __forceinline__ LdsFrames stage_frames(const DevP *__restrict__ tab, int nframes) {
    extern __shared__ __attribute__((aligned(16))) uint2 frames_lds[];
    constexpr int core_words = DEVP_CORE / (int)sizeof(uint2), full_words = (int)(sizeof(DevP) / sizeof(uint2));
    const uint2 *src = reinterpret_cast<const uint2 *>(tab);
    for (int i = threadIdx.x; i < nframes * core_words; i += blockDim.x) {
        const int f = i / core_words, w = i - f * core_words;
        frames_lds[i] = src[f * full_words + w];
    }
    __syncthreads();
    return LdsFrames{reinterpret_cast<const char *>(frames_lds)};
}
static_assert(sizeof(DevP) % sizeof(uint2) == 0, "DevP in 8-byte words");
constexpr int MAX_BATCH_FRAMES = 64;
static_assert(MAX_BATCH_FRAMES * DEVP_CORE <= 65536, "a batch's frame table fits one block's LDS");

// the frame of a batch that work item q belongs to
template <class TAB>
__device__ __forceinline__ int find_frame(const TAB &tab, int n, long long q) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (tab[mid].q_begin <= q)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

// The frame of a batch that a run of work items starts in, and the item
// range of that frame: a wave looks it up once per chunk it takes (a chunk
// of 256 items spans at most two frames of the usual sizes), and an item
// outside the range falls back to the search.
struct FrameHint {
    int fr;
    long long lo, hi;  // [lo, hi): the items of frame fr
};

template <class TAB>
__device__ __forceinline__ FrameHint frame_hint(const TAB &tab, int n, long long q) {
    const int f = find_frame(tab, n, q);
    return {f, tab[f].q_begin, f + 1 < n ? tab[f + 1].q_begin : (long long)1 << 62};
}

template <class TAB>
__device__ __forceinline__ int frame_of(const TAB &tab, int n, const FrameHint &h, long long q) {
    return q >= h.lo && q < h.hi ? h.fr : find_frame(tab, n, q);
}

// n / d and its remainder for 0 <= n < 2^31, 0 < d, without an integer
// divide: the float quotient, corrected to the exact one
__device__ __forceinline__ int div_rem(int n, int d, int &r) {
    int k = (int)((float)n * __builtin_amdgcn_rcpf((float)d));
    r = n - k * d;
    while (r < 0) {
        k--;
        r += d;
    }
    while (r >= d) {
        k++;
        r -= d;
    }
    return k;
}

// the pixel of item ql (counted from its frame's first item) of frame F's
// tile queue (false: the item is in a ragged edge tile, outside the
// picture); the tile sides are powers of two
__device__ __forceinline__ bool item_pixel(const DevP &F, int ql, int &x, int &yr) {
    const int tile = ql >> F.tn_shift;
    const int k = ql & ((1 << F.tn_shift) - 1);
    int tx;
    const int ty = div_rem(tile, F.tiles_x, tx);
    x = (tx << F.tw_shift) + (k & (F.tile_w - 1));
    yr = ty * F.tile_h + (k >> F.tw_shift);
    return x < F.width && yr < F.rows;
}

// The warm start's proposals: the network (BF16 MFMA, 64 rays a wave) on
// the ray of every work item of the queue, written to prop[q] (hit distance
// / 4, hit logit).  A kernel of its own, ahead of the march in the same
// stream: a march wave cannot take the proposals of a chunk another wave
// computed in the same kernel without an agent-scope release and acquire,
// which on this GPU write back and invalidate an XCD's whole L2 (measured:
// +0.75 ms a 1080p frame at a fence pair per chunk of 256 rays).
// TILES: the cone pass's gate instead -- the network on every tile's middle
// ray (the cone ray's), written to the slot of the tile's first item.
template <bool BATCH>
__global__ __launch_bounds__(256) void k_mlp_propose(DevP P0, const DevP *__restrict__ gtab, int nframes, MlpDev M,
                                                     const double2 *__restrict__ rtab, const ConeRes *__restrict__ cone,
                                                     float2 *prop, unsigned long long *__restrict__ ctr, int tiles) {
    const GlobalFrames tab{reinterpret_cast<const char *>(gtab)};
    __shared__ __attribute__((aligned(16))) char lds[4][MLP_LDS_BYTES];
    __shared__ __attribute__((aligned(16))) mlp_s4 wl[MLP_TILES * 64];
    mlp_stage_weights(M, wl);
    __syncthreads();
    const int lane = lane_id();
    // the units: work items, or tiles (each its first item's slot)
    const int ushift = tiles ? P0.tn_shift : 0;
    const long long units = P0.nslots_q >> ushift;
    const long long groups = (units + 63) / 64;
    const long long waves = (long long)gridDim.x * 4;
    unsigned long long evaluated = 0;  // rays the network ran on (the wave's, the same in every lane)
    // a uniform loop: every lane of a wave runs the same groups
    for (long long g = (long long)blockIdx.x * 4 + (threadIdx.x >> 6); g < groups; g += waves) {
        const long long u = g * 64 + lane;
        const long long q = u << ushift;
        float o[3] = {0.0f, 0.0f, 0.0f}, d[3] = {0.0f, 0.0f, 1.0f};
        int f = 0, x = 0, yr = 0;
        const bool in_q = u < units;
        if (BATCH) {
            const FrameHint h = frame_hint(tab, nframes, (g * 64) << ushift);  // wave-uniform
            f = in_q ? frame_of(tab, nframes, h, q) : 0;
        }
        const DevP &F = BATCH ? tab[f] : P0;
        if (tiles) {
            if (in_q) {
                int tx;
                const int ty = div_rem((int)((q - F.q_begin) >> F.tn_shift), F.tiles_x, tx);
                double cc[3], vv[3];
                frac_ray(F, (double)(tx << F.tw_shift) + (F.tile_w - 1.0) * 0.5,
                         (double)(F.row0 + ty * F.tile_h) + (F.tile_h - 1.0) * 0.5, cc, vv);
                const double il = 1.0 / sqrt(vv[0] * vv[0] + vv[1] * vv[1] + vv[2] * vv[2]);
                for (int j = 0; j < 3; j++) {
                    o[j] = (float)cc[j];
                    d[j] = (float)(vv[j] * il);
                }
            }
        } else {
            // the items of tiles the cone rays found empty take no proposal
            // (the march makes them background): a wave whose 64 items all
            // lie in such tiles skips the network
            const bool empty = !in_q || (cone && cone[q >> P0.tn_shift].t > F.zend);
            if (__all(empty)) {
                if (in_q) prop[q] = make_float2(0.0f, -1e30f);
                continue;
            }
            if (in_q && item_pixel(F, (int)(q - F.q_begin), x, yr)) {
                double cc[3], vv[3];
                pixel_ray_tab(F, rtab, x, yr, cc, vv);
                const double il = 1.0 / sqrt(vv[0] * vv[0] + vv[1] * vv[1] + vv[2] * vv[2]);
                for (int j = 0; j < 3; j++) {
                    o[j] = (float)cc[j];
                    d[j] = (float)(vv[j] * il);
                }
            }
        }
        // the network sees the ray as its point of closest approach to the
        // origin and its direction, and answers the hit's distance past that
        // point (a camera-independent form): back to the distance from the
        // start point
        const float s = -(o[0] * d[0] + o[1] * d[1] + o[2] * d[2]);
        const float pc[3] = {o[0] + s * d[0], o[1] + s * d[1], o[2] + s * d[2]};
        float t = 0.0f, lg = 0.0f;
        mlp_wave_infer(M, wl, pc, d, lds[threadIdx.x >> 6], t, lg);
        evaluated += 64;
        if (in_q) prop[q] = make_float2(t + 0.25f * s, lg);
    }
    if (lane_id() == 0 && evaluated) atomicAdd(&ctr[C_STATS + 15], evaluated);
}

// The cone rays (cone_march, src/calc.rs:2357-2435, and its grid,
// src/cone.rs:1-31), one lane per tile of the queue: the tile's coarse ray
// from its middle, marched in f64 from the start plane while the distance
// estimate keeps every ray of the tile -- within the spread a + slope t of
// the coarse ray (the widest gaps of the tile's starts and directions, in
// step widths), times the march's step factor margin -- clear of the DE
// stop.  Its depth is a safe start for the tile's rays; a cone that passes
// the end of the view makes the tile background (within the march's own
// assumption that the estimate is a bound).  One level of blocks: the
// queue's tile.
constexpr double kConeMinStep = 0.05, kConeMinShare = 0.1, kConeMargin = 1.5;  // src/calc.rs:2659-2674

// gate (not NULL): the network's logit per tile (k_mlp_propose's TILES
// pass); a tile at or above gate_logit marches no cone ray (depth 0: its
// rays march plainly).
// Persistent: one cone ray per lane, one f64 DE per turn, and a lane whose
// ray has ended takes the next tile of the queue (C_CONEQ, chunks of 64)
// once 8 of the wave's lanes are idle, so that a wave does not wait for its
// longest ray.
template <int KIND, int DEM, bool BATCH>
__global__ __launch_bounds__(256) void k_cone(DevP P0, const DevP *__restrict__ gtab, int nframes,
                                              const double2 *__restrict__ rtab, ConeRes *__restrict__ out,
                                              unsigned long long *__restrict__ ctr, const float2 *__restrict__ gate,
                                              float gate_logit) {
    const GlobalFrames tab{reinterpret_cast<const char *>(gtab)};
    Cnt cnt;
    const long long ntiles = P0.nslots_q >> P0.tn_shift;
    WaveChunk tq;
    chunk_init(tq);
    bool have = false, exhausted = false;
    long long g = 0;  // the lane's tile
    int fr = 0;
    double c[3] = {0, 0, 0}, v0[3] = {0, 0, 0};
    double a = 0.0, slope = 0.0, lip = 0.0, div = 1.0;
    double last_de = 0.0, last_step = 0.0, t = 0.0;
    float count = 0.0f;
    uint32_t guard = 0;
    for (;;) {
        bool need = !have && !exhausted;
        const int idle = __popcll(__ballot(need));
        const bool refill = idle > 0 && (idle >= 8 || idle == __popcll(__ballot(!exhausted)));
        for (int tries = 0; refill && tries < 4 && __any(need); tries++) {
            const long long gi = wave_take(need, &ctr[C_CONEQ], tq, 64, ntiles);
            if (!need) continue;
            if (gi < 0) {
                exhausted = true;
                need = false;
                continue;
            }
            const long long q = gi << P0.tn_shift;
            if (gate && gate[q].y >= gate_logit) {
                out[gi] = ConeRes{0.0, 0.0f, 0};
                continue;
            }
            need = false;
            have = true;
            g = gi;
            fr = BATCH ? find_frame(tab, nframes, q) : 0;
            const DevP &P = BATCH ? tab[fr] : P0;
            int tx;
            const int ty = div_rem((int)((q - P.q_begin) >> P.tn_shift), P.tiles_x, tx);
            const int x0 = tx << P.tw_shift, y0 = ty * P.tile_h;  // (y0: a row of the rectangle)
            const double fy0 = (double)(P.row0 + y0);
            double c0[3];
            frac_ray(P, (double)x0 + (P.tile_w - 1.0) * 0.5, fy0 + (P.tile_h - 1.0) * 0.5, c0, v0);
            // the tile's spread from the coarse ray, at the start and per
            // step width of depth (step widths): the widest gaps are the
            // corners' (the starts are affine in the pixel, and the
            // directions' angle from the middle one grows with both
            // offsets), so the four corner rays give the maxima over the
            // tile's rays
            a = 0.0;
            slope = 0.0;
            const int sj = P.tile_h > 1 ? P.tile_h - 1 : 1, si = P.tile_w > 1 ? P.tile_w - 1 : 1;
            for (int j = 0; j < P.tile_h; j += sj) {
                for (int i = 0; i < P.tile_w; i += si) {
                    double cc[3], vv[3];
                    if (x0 + i < P.width && y0 + j < P.rows)
                        pixel_ray_tab(P, rtab, x0 + i, y0 + j, cc, vv);
                    else
                        frac_ray(P, (double)(x0 + i), fy0 + j, cc, vv);
                    const double dc[3] = {cc[0] - c0[0], cc[1] - c0[1], cc[2] - c0[2]};
                    const double dv[3] = {vv[0] - v0[0], vv[1] - v0[1], vv[2] - v0[2]};
                    a = fmax(a, sqrt(dc[0] * dc[0] + dc[1] * dc[1] + dc[2] * dc[2]));
                    slope = fmax(slope, sqrt(dv[0] * dv[0] + dv[1] * dv[1] + dv[2] * dv[2]));
                }
            }
            a /= P.step_width;
            slope /= P.step_width;
            lip = kConeMargin / (double)fminf(fmaxf(P.z_step_div, 0.05f), 1.0f);
            div = 1.0 + slope * lip + (double)P.de_stop * (double)P.de_stop_factor;
            last_de = __builtin_inf();
            last_step = 0.0;
            t = 0.0;
            for (int k = 0; k < 3; k++) c[k] = c0[k];
            count = 0.0f;
            guard = 0;
            cnt.cone_rays++;
        }
        if (__all(exhausted && !have)) break;
        if (!have) continue;
        const DevP &P = BATCH ? tab[fr] : P0;
        // one step of the cone ray
        bool done = ++guard > 2000000u;
        const float ms = (float)((double)P.de_stop * (1.0 + t * (double)P.de_stop_factor));
        const De64 r = de64<KIND, DEM, true>(P, c[0], c[1], c[2], P.max_it, 0, ms);
        cnt.de64++;
        cnt.p8_64 += r.p8;
        cnt.box_64 += r.box;
        if (!done && r.it_result >= r.max_its) done = true;
        if (!done) {
            double de = r.de;
            if (de > last_de + last_step) de = last_de + last_step;
            const double rsf =
                last_step > 0.0 && last_de > de + 1e-30 ? fmin(fmax(last_step / (last_de - de), 0.5), 1.0) : 1.0;
            const double stop = (double)ms;
            const double reach = (a + slope * t) * lip + stop;
            const double cap = (double)(fmaxf(ms, 0.4f) * P.mh04zsd);
            const double sd = fmin((de - reach - (double)P.ms_de_sub * stop) * (double)P.z_step_div * rsf / div, cap);
            const double d = fmax(0.11, (de - (double)P.ms_de_sub * stop) * (double)P.z_step_div * rsf);
            if (sd < kConeMinStep * stop || sd < kConeMinShare * fmin(d, cap)) {
                done = true;
            } else {
                if (P.dfog_on_it == 0 || r.it_result == P.dfog_on_it) count += (float)(fmin(sd, P.zend - t) / d);
                if (t + sd > P.zend) {
                    t = __builtin_inf();
                    done = true;
                } else {
                    t += sd;
                    for (int k = 0; k < 3; k++) c[k] += v0[k] * sd;
                    last_de = de;
                    last_step = sd;
                }
            }
        }
        if (done) {
            out[g] = ConeRes{t, count, 0};
            have = false;
        }
    }
    cnt.cone_its = cnt.p8_64 + cnt.box_64 + cnt.oth_64;
    cnt.cone_des = cnt.de64;
    flush_counters(cnt, ctr);
}

// The tiles whose cone ray passed the end of the view, each by one whole
// wave: every pixel's start DE (march_pixel_from's check for a start in the
// set, src/calc.rs:2263-2290, as k_march32 makes it), then the background
// with the cone's fog count (src/calc.rs:2299-2305), written by the wave in
// whole rows (of four 8 x 8 tiles side by side where they are background
// together), or the pixel handed to the f64 finish when its start is in the
// set.  Every other tile is listed for the march (ConeRes::live,
// C_LIVE), so the march's persistent waves refill their lanes only with rays
// that march.  A wave takes 64 tiles at a time, lane l deciding tile g0 + l.
template <int KIND, int DEM, bool BATCH>
__global__ __launch_bounds__(256) void k_cone_bg(DevP P0, Bufs B, const DevP *__restrict__ gtab, int nframes) {
    const LdsFrames tab = BATCH ? stage_frames(gtab, nframes) : LdsFrames{nullptr};
    Cnt cnt;
    const long long ntiles = P0.nslots_q >> P0.tn_shift;
    const int tsize = 1 << P0.tn_shift;
    const long long wave = ((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const long long nwaves = ((long long)gridDim.x * blockDim.x) >> 6;
    const float fguard = P0.guard;
    for (long long g0 = wave * 64; g0 < ntiles; g0 += nwaves * 64) {
        const long long gi = g0 + lane_id();
        bool bg = false, live = false;
        int fr = 0;
        float count = 0.0f;
        if (gi < ntiles) {
            const ConeRes cr = B.cone[gi];
            fr = BATCH ? find_frame(tab, nframes, gi << P0.tn_shift) : 0;
            const DevP &F = BATCH ? tab[fr] : P0;
            bg = cr.t > F.zend;
            live = !bg;
            count = cr.count;
        }
        const long long s = wave_append(live, &B.ctr[C_LIVE]);
        if (live) B.cone[s].live = (int32_t)((uint32_t)gi | ((uint32_t)fr << LIVE_TILE_BITS));
        // one pixel of a background tile per lane (item q of frame tfr; `in`:
        // the lane has one): the start DE, then the background or a hand-off.
        // Called with the whole wave.
        auto bg_pixel = [&](int tfr, long long q, bool in, float tcount) {
            const DevP &F = BATCH ? tab[tfr] : P0;
            bool handoff = false, first = false;
            int x = 0, yr = 0, pix = 0;
            int32_t seed = 0;
            double c[3] = {0, 0, 0}, v[3] = {0, 0, 0};
            if (in && item_pixel(F, (int)(q - F.q_begin), x, yr)) {
                pix = (int)(F.pix_base + (long long)yr * F.width + x);
                // the start point (the direction only for a hand-off)
                pixel_start(F, x, yr, c);
                int it = 0;
                const float ms = F.de_stop;
                const float d32 = de32<KIND, DEM>(F, c, ms, it, cnt);
                if (!(d32 < __builtin_inff())) {  // single precision overflow: f64 decides
                    cnt.overflow++;
                    handoff = true;
                } else if (it >= F.max_it || !(d32 >= ms * fguard)) {
                    handoff = true;
                } else {
                    cnt.cone_bg++;
                    write_background(F, B, pix, tcount);
                }
            }
            // (an exact append: these hand-offs are rare, and the list's
            // capacity allows empty chunk tails for the march's waves only)
            const long long slot = wave_append(handoff, &B.ctr[C_CAND]);
            if (handoff) {
                pixel_ray_tab(F, B.rtab, x, yr, c, v);
                first = F.first_step_random != 0;
                seed = first ? dev_pixel_seed(F.row0 + yr, x) : 0;
                Cand &o = B.cand[slot];
                for (int i = 0; i < 3; i++) {
                    o.c[i] = c[i];
                    o.v[i] = v[i];
                }
                o.mzz = 0.0;
                o.last_step = 0.0;
                o.last_de = 0.0;
                o.rsf = 1.0f;
                o.step_count = 0.0f;
                o.pix = pix;
                o.flags = CF_START | (first ? CF_FIRST : 0) | (tfr << 8);
                o.nsteps = 0;
                o.seed = seed;
            }
        };
        const unsigned long long bgm = __ballot(bg);
        // Quads of four background 8 x 8 tiles side by side, the first at a
        // multiple of 4 in its row (rows of a multiple of 4 tiles hold whole
        // quads, frames whole rows): the wave takes them as rows of 32
        // pixels, so that its depth and normal writes fill whole 128-byte
        // lines.  Lane l of pass p: row 2p + l / 32, column l % 32.
        unsigned long long quads = 0;
        if (P0.tile_w == 8 && P0.tile_h == 8) {
            bool head = false;
            if ((lane_id() & 3) == 0 && ((bgm >> lane_id()) & 15ull) == 15ull) {
                const DevP &F = BATCH ? tab[fr] : P0;
                const long long tf = gi - (F.q_begin >> P0.tn_shift);  // the tile's index in its frame
                head = (F.tiles_x & 3) == 0 && (tf & 3) == 0;
            }
            quads = __ballot(head);
        }
        for (unsigned long long qm = quads; qm; qm &= qm - 1) {
            const int l = __ffsll((long long)qm) - 1;
            const int tfr = __shfl(fr, l, 64);
            const int col = lane_id() & 31, t = col >> 3;
            const float tcount = __shfl(count, l + t, 64);
            for (int p = 0; p < 4; p++) {
                const int r = 2 * p + (lane_id() >> 5);
                bg_pixel(tfr, ((g0 + l + t) << P0.tn_shift) + r * 8 + (col & 7), true, tcount);
            }
        }
        // the other background tiles one at a time: lane l takes item l
        unsigned long long m = bgm & ~(quads | quads << 1 | quads << 2 | quads << 3);
        while (m) {
            const int l = __ffsll((long long)m) - 1;
            m &= m - 1;
            const long long tq = (g0 + l) << P0.tn_shift;  // the tile's first item
            const int tfr = __shfl(fr, l, 64);
            const float tcount = __shfl(count, l, 64);
            for (int k0 = 0; k0 < tsize; k0 += 64) bg_pixel(tfr, tq + k0 + lane_id(), k0 + lane_id() < tsize, tcount);
        }
    }
    flush_counters(cnt, B.ctr);
}

// BATCH: the frames of a batch in one launch; `tab` holds each frame's
// parameters, P0 the batch's queue totals and the shared tile shape.
// LEARN: the warm start (B.prop holds k_mlp_propose's proposals).  A ray
// starts where the network puts it: confident hits at the predicted
// distance less a margin, confident misses past the end (the background).
// A start that is not free -- the DE there below the stop distance (with
// the hand-off guard) or the iteration limit reached -- is rolled back to
// the plain march from the start plane.  This is the reference's cone-march
// jump and check (march_pixel_from, src/calc.rs:2293-2308) with the
// network's depth in place of the coarse ray's.
template <int KIND, int DEM, bool BATCH, bool LEARN>
__global__ __launch_bounds__(256) void k_march32(DevP P0, Bufs B, const DevP *__restrict__ gtab, int nframes,
                                                 MlpDev M) {
    const LdsFrames tab = BATCH ? stage_frames(gtab, nframes) : LdsFrames{nullptr};
    Cnt cnt;
    int fr = 0;
    bool have = false, exhausted = false;
    int pix = 0, px = 0, py = 0;
    double c[3] = {0, 0, 0}, v[3] = {0, 0, 0};
    double mzz = 0;
    float last_step = 0, last_de = 0, dtmp = 0;
    float ms = 0, rsf = 1, scount = 0;
    int nsteps = 0;
    bool start = false, first = false, jumped = false;
    float jrem = 0.0f;  // a jump's predicted distance left to the surface (march units)
    double cone_t = 0.0;  // the lane's tile's cone depth (0: none)
    float cone_cnt = 0.0f;
    int32_t seed = 0;
    const float fguard = P0.guard;
    WaveChunk wq, cq;
    chunk_init(wq);
    chunk_init(cq);
    FrameHint hint = {0, 0, 0};
    // the queue: every tile's items, or with the cone rays the items of the
    // tiles k_cone_bg listed (the others are background or handed off)
    const bool listed = P0.cone_mode != 0;
    const long long n_items = listed ? (long long)B.ctr[C_LIVE] << P0.tn_shift : P0.nslots_q;
    auto on_chunk = [&](long long base, int) {
        if (BATCH && !listed) hint = frame_hint(tab, nframes, base);
    };
    for (;;) {
        // refill: the idle lanes take the next work items from the wave's
        // chunk, once at least refill_min of them are idle (or every lane
        // that still has work to get is)
        bool need = !have && !exhausted;
        const int idle = __popcll(__ballot(need));
        const bool refill = idle > 0 && (idle >= P0.refill_min || idle == __popcll(__ballot(!exhausted)));
        for (int tries = 0; refill && tries < 4 && __any(need); tries++) {
            const long long qi = wave_take(need, &B.ctr[C_QUEUE], wq, 256, n_items, on_chunk);
            if (need) {
                int x = 0, yr = 0;
                if (qi < 0) {
                    exhausted = true;
                    need = false;
                    continue;
                }
                long long q = qi;  // the item of the tile queue
                if (listed) {
                    const uint32_t e = (uint32_t)B.cone[qi >> P0.tn_shift].live;
                    q = ((long long)(e & ((1u << LIVE_TILE_BITS) - 1u)) << P0.tn_shift) |
                        (qi & ((1 << P0.tn_shift) - 1));
                    fr = BATCH ? (int)(e >> LIVE_TILE_BITS) : 0;
                } else {
                    fr = BATCH ? frame_of(tab, nframes, hint, q) : 0;
                }
                const DevP &F = BATCH ? tab[fr] : P0;
                if (item_pixel(F, (int)(q - F.q_begin), x, yr)) {
                    have = true;
                    need = false;
                    px = x;
                    py = yr;  // the row within the frame's rectangle
                    pix = (int)(F.pix_base + (long long)yr * F.width + x);
                    pixel_ray_tab(F, B.rtab, px, py, c, v);
                    mzz = 0.0;
                    ms = F.de_stop;
                    rsf = 1.0f;
                    scount = 0.0f;
                    nsteps = 0;
                    start = true;
                    jumped = false;
                    first = F.first_step_random != 0;
                    seed = first ? dev_pixel_seed(F.row0 + py, px) : 0;
                    cone_t = 0.0;
                    if (P0.cone_mode) {
                        const ConeRes cr = B.cone[q >> P0.tn_shift];
                        cone_t = cr.t;
                        cone_cnt = cr.count;
                    }
                    // (a tile its cone ray found empty takes no network jump:
                    // its rays' first DE is at the start plane)
                    if (LEARN && !(cone_t > F.zend)) {
                        const float2 pr = B.prop[q];
                        const float ph = 1.0f / (1.0f + __expf(-pr.y));
                        double jump = 0.0;
                        if (ph >= M.conf_hit) {
                            const double vl = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
                            const double th = 4.0 * (double)pr.x;
                            const double tj = th * (1.0 - (double)M.margin_rel) - M.margin_abs;
                            if (tj > 0.0) {
                                jump = tj / vl;
                                jrem = (float)((th - tj) / vl);
                            }
                        } else if (ph <= M.conf_miss) {
                            jump = F.zend * 2.0;
                        }
                        if (jump > F.zend) {  // past the end: the background
                            write_background(F, B, pix, 0.0f);
                            cnt.jumps++;
                            have = false;
                            need = true;
                        } else if (jump > 0.0) {
                            mzz = jump;
                            c[0] += v[0] * jump;
                            c[1] += v[1] * jump;
                            c[2] += v[2] * jump;
                            ms = F.de_stop * (1.0f + (float)mzz * F.de_stop_factor);
                            jumped = true;
                        }
                    }
                }
            }
        }
        const DevP &P = BATCH ? tab[fr] : P0;
        if (__all(exhausted)) break;
        bool handoff = false, background = false;
        if (have) {
            int it = 0;
            const float d32 = de32<KIND, DEM>(P, c, ms, it, cnt);
            // the step bookkeeping of the FP32 march runs in f32 (the f64
            // finish re-decides everything near the surface); the depth and
            // the position accumulate in f64
            bool top = true;
            if (jumped) {
                // the warm start's check (a network jump, or a jump to the
                // cone's depth): a free start, clear of the hand-off guard and
                // of a tenth of the distance the network left to the surface,
                // is kept (a DE far under that distance means a surface
                // nearer than predicted: one in front, or one the jump passed
                // through); any other goes back to the start plane and marches
                // plainly (as march_pixel does when the cone's jump lands in
                // the set, src/calc.rs:2207-2229, 2305-2319)
                jumped = false;
                if (!(d32 < __builtin_inff()) || it >= P.max_it || !(d32 >= ms * fguard) ||
                    !(d32 >= 0.1f * jrem)) {
                    cnt.rollbacks++;
                    pixel_ray_tab(P, B.rtab, px, py, c, v);
                    mzz = 0.0;
                    ms = P.de_stop;
                    scount = 0.0f;
                    cone_t = 0.0;
                    top = false;
                } else {
                    cnt.jumps++;
                }
            }
            if (!top) {
            } else if (!(d32 < __builtin_inff())) {  // single precision overflow: f64 decides
                cnt.overflow++;
                handoff = true;
                top = false;
            } else if (start) {
                if (it >= P.max_it || !(d32 >= ms * fguard)) {
                    handoff = true;
                    top = false;
                } else if (cone_t > P.zend) {
                    // the tile's cone ray passed the end of the view: the
                    // background, with the dynamic fog's count of the way
                    // (march_pixel_from, src/calc.rs:2299-2305)
                    background = true;
                    scount = cone_cnt;
                    cnt.cone_bg++;
                    top = false;
                } else if (P.cone_mode == 2 && cone_t > mzz) {
                    // the cone's depth: the march goes on from there, checked
                    // as a jump (src/calc.rs:2306-2318)
                    const double dj = cone_t - mzz;
                    mzz = cone_t;
                    c[0] += v[0] * dj;
                    c[1] += v[1] * dj;
                    c[2] += v[2] * dj;
                    ms = P.de_stop * (1.0f + (float)mzz * P.de_stop_factor);
                    scount = cone_cnt;
                    jumped = true;
                    jrem = 0.0f;
                    cone_t = 0.0;
                    top = false;
                } else {
                    dtmp = d32;
                    last_step = dtmp * P.z_step_div;
                }
            } else {
                dtmp = fminf(d32, last_de + last_step);
                if (last_de > dtmp + 1e-30f) {
                    const float t = last_step * __builtin_amdgcn_rcpf(last_de - dtmp);
                    rsf = t < 1.0f ? fmaxf(0.5f, t) : 1.0f;
                } else {
                    rsf = 1.0f;
                }
            }
            if (top) {
                if (it >= P.max_it) {
                    handoff = true;  // the half step back (src/calc.rs:2545-2552) is f64 work
                } else if (it < P.min_it || dtmp >= ms * fguard) {
                    if (P.step_cap > 0 && nsteps >= P.step_cap) {
                        background = true;
                    } else if (P.it_split > 0 && nsteps >= P.it_split) {
                        // the iteration-cap split: a ray still marching after
                        // it_split steps goes on in the f64 finish, which
                        // resumes it from here, so that the long rays of a
                        // launch do not hold the FP32 pass's tail
                        handoff = true;
                    } else {
                        last_de = dtmp;
                        float d = fmaxf(0.11f, (dtmp - P.ms_de_sub * ms) * P.z_step_div * rsf);
                        const float s1 = fmaxf(ms, 0.4f) * P.mh04zsd;
                        const bool count = P.dfog_on_it == 0 || it == P.dfog_on_it;
                        if (s1 < d) {
                            if (count) scount += s1 * __builtin_amdgcn_rcpf(d);
                            d = s1;
                        } else if (count) {
                            scount += 1.0f;
                        }
                        if (first) {
                            first = false;
                            seed = (int32_t)((uint32_t)seed * 214013u + 2531011u);
                            d = (float)((double)d * ((double)(seed & 0x7FFFFFFF) * (1.0 / (double)0x7FFFFFFF)));
                        }
                        const double dd = (double)d;
                        mzz += dd;
                        if (mzz > P.zend) {
                            background = true;
                        } else {
                            nsteps++;
                            cnt.steps++;
                            last_step = d;
                            c[0] += v[0] * dd;
                            c[1] += v[1] * dd;
                            c[2] += v[2] * dd;
                            ms = P.de_stop * (1.0f + (float)mzz * P.de_stop_factor);
                            start = false;
                        }
                    }
                } else {
                    handoff = true;
                }
            }
        }
        const long long slot = wave_take(handoff, &B.ctr[C_CAND], cq, 64, -1);
        if (handoff) {
            Cand &o = B.cand[slot];
            o.c[0] = c[0];
            o.c[1] = c[1];
            o.c[2] = c[2];
            o.v[0] = v[0];
            o.v[1] = v[1];
            o.v[2] = v[2];
            o.mzz = mzz;
            o.last_step = last_step;
            o.last_de = last_de;
            o.rsf = rsf;
            o.step_count = scount;
            o.pix = pix;
            o.flags = (start ? CF_START : 0) | (first ? CF_FIRST : 0) | (nsteps > 0 ? CF_STEPPED : 0) | (fr << 8);
            o.nsteps = nsteps;
            o.seed = seed;
            have = false;
        }
        if (background) {
            write_background(P, B, pix, scount);
            have = false;
        }
    }
    // the wave's unused hand-off slots are marked empty
    if (lane_id() < cq.left) B.cand[cq.base + lane_id()].pix = -1;
    flush_counters(cnt, B.ctr);
}

// ---------------------------------------------------------------------------
// pass 2: the f64 finish.  Each lane is a resumable state machine with one
// DE site (the reference's WGSL keeps the same shape, one evaluate() call in
// a loop of states, src/gpu_march.wgsl:1903): lanes in different states --
// march, bisection, normals -- share the one f64 DE of every turn, and a lane
// that finishes takes the next ray from the list at once.

// store_normal (src/calc.rs:1794-1804)
__device__ inline void store_normal(const double n[3], int16_t out[3]) {
    const double d = 32767.0 / sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2] + 1e-100);
    for (int k = 0; k < 3; k++) {
        const double v = rte(n[k] * d);
        out[k] = isnan(v) ? 0 : (int16_t)fmin(fmax(v, -32767.0), 32767.0);
    }
    if (out[0] == 0 && out[1] == 0) out[2] = out[2] > 0 ? 32767 : -32767;
}

// the surface values of surface_found (src/calc.rs:2471-2517) into the outputs
__device__ inline void write_surface(const DevP &P, const Bufs &B, int pix, const Ray64 &R, const double n[3], float nn,
                              bool de_limited, uint16_t iters, float step_count) {
    if (B.depth) B.depth[pix] = (float)R.mzz;
    if (B.depth64) B.depth64[pix] = R.mzz;
    if (B.iters) B.iters[pix] = iters;
    if (B.normals) {
        const double l = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
        for (int k = 0; k < 3; k++) B.normals[3 * pix + k] = l > 0 ? (float)(n[k] / l) : 0.0f;
    }
    // the G-buffer's words (their f64 square root, divide and logarithm
    // only when a G-buffer is written)
    if (B.gbuf) {
        int16_t nq[3];
        store_normal(n, nq);
        float g;
        if (de_limited) {
            const float lg = (float)log((double)fmaxf(P.de_stop, R.ms) * P.step_width);
            g = 32767.0f - (nn + P.d_col_plus + P.col_var * lg) * P.mcts_m;
        } else {
            g = 32767.0f - nn * P.mcts_m;
        }
        const uint32_t si_gradient = clip15(g);
        const uint32_t otrap = clip15((float)(R.otrap * 4096.0));  // do_color, option 0
        const uint32_t sh = (uint32_t)rtef(fminf(fmaxf(step_count, 0.0f), 1023.0f));
        uint32_t *w = B.gbuf + 5 * (size_t)pix;
        w[0] = (uint32_t)(uint16_t)nq[0] | ((uint32_t)(uint16_t)nq[1] << 16);
        w[1] = (uint32_t)(uint16_t)nq[2] | (sh << 16);
        w[2] = R.zpos_fine(P);
        w[3] = si_gradient | (otrap << 16);
        w[4] = 5000;
    }
}

// a ray that starts inside the set (src/calc.rs:2277-2293)
__device__ inline void write_inside(const DevP &P, const Bufs &B, int pix, const Ray64 &R) {
    const double t = fmin(fmax(R.rout / P.d_rstop, 0.0), 1.0);
    const uint32_t sg = 32768u + (uint32_t)rte(32767.0 * t);
    const uint32_t otrap = clip15((float)(R.otrap * 4096.0));
    if (B.depth) B.depth[pix] = 0.0f;
    if (B.depth64) B.depth64[pix] = 0.0;
    if (B.iters) B.iters[pix] = (uint16_t)max(R.it, 0);
    if (B.normals) {
        B.normals[3 * pix] = 0.0f;
        B.normals[3 * pix + 1] = 0.0f;
        B.normals[3 * pix + 2] = -1.0f;
    }
    if (B.gbuf) {
        uint32_t *w = B.gbuf + 5 * (size_t)pix;
        w[0] = 0;
        w[1] = (uint32_t)(uint16_t)(int16_t)-32767;
        w[2] = 0x7FFF0000u;
        w[3] = sg | (otrap << 16);
        w[4] = 5000;
    }
}

enum FinPhase : int { FP_START, FP_RESUME, FP_BACK, FP_STEP, FP_REF };

// The finish ends at the refined surface point: its record goes to the
// surface list, and k_normals computes the normal, colour and G-buffer
// words.  So that every phase here makes the same f64 DE: one DE site per
// turn.
// The batches' finish (BATCH and FAST) is built for four waves per SIMD
// (128 VGPRs, 16 of them spilled): at the compiler's three (155 VGPRs) its
// waves waited a third of their cycles, and four hide more of that latency
// (6.60 -> 6.24 ms a 48-frame launch).  A single frame's finish spills more
// at four (41 VGPRs) and was slower there (0.84 -> 1.15 ms with the normals).
template <int KIND, int DEM, bool BATCH, bool FAST>
__global__ __launch_bounds__(256, (BATCH && FAST) ? 4 : 1) void k_finish64(DevP P0, Bufs B, int all_pixels,
                                                                           const DevP *__restrict__ gtab,
                                                                           int nframes) {
    const LdsFrames tab = BATCH ? stage_frames(gtab, nframes) : LdsFrames{nullptr};
    Cnt cnt;
    int fr = 0;
    const long long n_items = all_pixels ? P0.npix : (long long)B.ctr[C_CAND];
    // the list mode's surface records: one per hand-off slot
    if (!all_pixels && blockIdx.x == 0 && threadIdx.x == 0) B.ctr[C_SURF] = (unsigned long long)n_items;
    long long cur = 0;  // the lane's hand-off slot
    bool have = false, exhausted = false;
    Ray64 R;
    Refine F;
    int pix = 0, phase = FP_START;
    double last_step = 0, last_de = 0, dtmp = 0, back_dt1 = 0;
    float rsf = 1.0f, scount = 0.0f;
    int nsteps = 0;
    uint32_t guard = 0;
    bool first = false, stepped = false, de_limited = false;
    int32_t seed = 0;
    WaveChunk fq, sq;
    chunk_init(fq);
    chunk_init(sq);
    for (;;) {
        // refill as the FP32 march does: once refill_min lanes are idle, or
        // every lane that still has work to get is
        bool need = !have && !exhausted;
        const int idle = __popcll(__ballot(need));
        const bool refill = idle > 0 && (idle >= P0.refill_min || idle == __popcll(__ballot(!exhausted)));
        for (int tries = 0; refill && tries < 4 && __any(need); tries++) {
            const long long slot = wave_take(need, &B.ctr[C_FINQ], fq, 64, n_items);
            if (!need) continue;
            if (slot < 0) {
                exhausted = true;
                need = false;
                continue;
            }
            if (!all_pixels && B.cand[slot].pix < 0) {  // an empty slot of a wave's chunk
                B.surf[slot].pix = -1;
                continue;
            }
            have = true;
            need = false;
            cur = slot;
            if (!all_pixels && BATCH) fr = B.cand[slot].flags >> 8;
            const DevP &F = BATCH ? tab[fr] : P0;
            R.reset(F);
            guard = 0;
            if (all_pixels) {
                pix = (int)slot;
                const int x = pix % F.width, y = F.row0 + pix / F.width;
                pixel_ray(F, x, y, R.c, R.v);
                phase = FP_START;
                first = F.first_step_random != 0;
                seed = first ? dev_pixel_seed(y, x) : 0;
                scount = 0.0f;
                nsteps = 0;
                rsf = 1.0f;
                stepped = false;
            } else {
                const Cand &o = B.cand[slot];
                pix = o.pix;
                for (int k = 0; k < 3; k++) {
                    R.c[k] = o.c[k];
                    R.v[k] = o.v[k];
                }
                last_step = o.last_step;
                last_de = o.last_de;
                rsf = o.rsf;
                scount = o.step_count;
                nsteps = o.nsteps;
                seed = o.seed;
                first = (o.flags & CF_FIRST) != 0;
                stepped = (o.flags & CF_STEPPED) != 0;
                if (o.flags & CF_START) {
                    phase = FP_START;
                } else {
                    R.mzz = o.mzz;
                    R.upd(F);
                    phase = FP_RESUME;
                }
            }
        }
        if (__all(exhausted)) break;
        const DevP &P = BATCH ? tab[fr] : P0;
        bool surf = false, done = false;
        if (!have) goto end;
        R.take(de64<KIND, DEM, FAST>(P, R.c[0], R.c[1], R.c[2], R.max_it, R.calc_sit, R.ms), cnt);
        switch (phase) {
            case FP_START:
                if (R.it >= R.max_its || R.de < (double)R.ms) {
                    write_inside(P, B, pix, R);
                    done = true;
                    goto end;
                }
                dtmp = R.de;
                last_step = dtmp * (double)P.z_step_div;
                goto top;
            case FP_RESUME:
                dtmp = R.de;
                if (!stepped) {
                    last_step = dtmp * (double)P.z_step_div;
                    goto top;
                }
                goto clamp;
            case FP_BACK:
                dtmp = R.de;
                last_step = -back_dt1;
                goto cond;
            case FP_STEP:
                dtmp = R.de;
                goto clamp;
            default:  // FP_REF
                if (F.after(R, P)) goto end;
                goto refined;
        }
    clamp:
        // after a step of march_from (src/calc.rs:2626-2638)
        if (dtmp > last_de + last_step) dtmp = last_de + last_step;
        if (last_de > dtmp + 1e-30) {
            const double t = last_step / (last_de - dtmp);
            rsf = t < 1.0 ? (float)fmax(0.5, t) : 1.0f;
        } else {
            rsf = 1.0f;
        }
    top:
        // the head of march_from's loop (src/calc.rs:2540-2552)
        if (++guard > 2000000u) {
            write_background(P, B, pix, scount);
            done = true;
            goto end;
        }
        if (R.it >= R.max_its) {
            back_dt1 = -0.5 * last_step;
            R.move(back_dt1);
            R.upd(P);
            phase = FP_BACK;
            goto end;
        }
    cond:
        if (R.it < P.min_it || (R.it < R.max_its && dtmp >= (double)R.ms)) {
            // the next step (src/calc.rs:2553-2620)
            if (P.step_cap > 0 && nsteps >= P.step_cap) {
                write_background(P, B, pix, scount);
                done = true;
                goto end;
            }
            last_de = dtmp;
            double d = fmax(0.11, (dtmp - (double)P.ms_de_sub * (double)R.ms) * (double)P.z_step_div * (double)rsf);
            const double s1 = (double)(fmaxf(R.ms, 0.4f) * P.mh04zsd);
            const bool count = P.dfog_on_it == 0 || R.it == P.dfog_on_it;
            if (s1 < d) {
                if (count) scount += (float)(s1 / d);
                d = s1;
            } else if (count) {
                scount += 1.0f;
            }
            if (first) {
                first = false;
                seed = (int32_t)((uint32_t)seed * 214013u + 2531011u);
                d *= (double)(seed & 0x7FFFFFFF) * (1.0 / (double)0x7FFFFFFF);
            }
            R.mzz += d;
            if (R.mzz > P.zend) {
                write_background(P, B, pix, scount);
                done = true;
                goto end;
            }
            nsteps++;
            cnt.steps++;
            cnt.steps64++;
            last_step = d;
            for (int k = 0; k < 3; k++) R.c[k] += R.v[k] * d;
            R.upd(P);
            phase = FP_STEP;
            goto end;
        }
        // surface found (src/calc.rs:2471-2481)
        de_limited = R.it < R.max_its || dtmp < (double)R.ms;
        if (P.de_add_steps != 0) {
            const bool more = de_limited ? F.start_bs(R, P, dtmp, last_step, P.de_add_steps)
                                         : F.start_bsi(R, P.de_add_steps);
            if (more) {
                phase = FP_REF;
                goto end;
            }
        }
    refined:
        // the refined surface point (surface_found, src/calc.rs:2471-2481):
        // its normal and colour follow in k_normals
        surf = true;
        done = true;
    end:
        if (done) have = false;
        // the surface record: in the list mode at the ray's hand-off slot
        // (an empty record where the ray ended without a surface), in the
        // all-pixels mode appended
        long long ss = cur;
        if (all_pixels) ss = wave_take(surf, &B.ctr[C_SURF], sq, 64, -1);
        if (done && !surf && !all_pixels) B.surf[cur].pix = -1;
        if (surf) {
            SurfRec &o = B.surf[ss];
            for (int k = 0; k < 3; k++) o.c[k] = R.c[k];
            o.mzz = R.mzz;
            o.ms = R.ms;
            o.step_count = scount;
            o.pix = pix;
            o.info = min(max(R.it, 0), (int)SR_ITERS) | (de_limited ? SR_DELIM : 0) | (fr << 20);
        }
    }
    if (lane_id() < sq.left) B.surf[sq.base + lane_id()].pix = -1;
    flush_counters(cnt, B.ctr);
}

// The normals pass over the refined surface points: calculate_normals with
// smooth_normals 0 (src/calc.rs:1813-1858) -- the centre DE with the smooth
// iterations, then the DEs at +-noffset along vg2, vg0, vg1 -- and the
// surface's G-buffer words (write_surface).  Every lane makes the same seven
// DEs, so the waves stay converged.  N32: the six offset DEs in FP32, for
// the analytic DEs when the stencil's offset spans enough FP32 ulps of the
// coordinates (the host's fp32_stencil_ok; at the 1080p Mandelbulb their
// error moves a normal by ~0.01 degree); the numeric gradient's DE does not
// allow it.
template <int KIND, int DEM, bool N32, bool BATCH>
__global__ __launch_bounds__(256) void k_normals(DevP P0, Bufs B, const DevP *__restrict__ gtab, int nframes) {
    const LdsFrames tab = BATCH ? stage_frames(gtab, nframes) : LdsFrames{nullptr};
    Cnt cnt;
    const long long n = (long long)B.ctr[C_SURF];
    const long long stride = (long long)gridDim.x * blockDim.x;
    const SurfRec none{{0, 0, 0}, 0, 0, 0, -1, 0};
    // whole waves step together; each turn loads the next turn's record
    // ahead of its DEs (the load's latency hidden behind them)
    long long base = (long long)blockIdx.x * blockDim.x + (threadIdx.x & ~63);
    SurfRec next = base + lane_id() < n ? B.surf[base + lane_id()] : none;
    for (; base < n; base += stride) {
        const long long i = base + lane_id();
        const SurfRec o = next;
        next = i + stride < n ? B.surf[i + stride] : none;
        bool hit = false;
        int pix = 0;
        if (o.pix >= 0) {
            const int fr = BATCH ? o.info >> 20 : 0;
            const DevP &P = BATCH ? tab[fr] : P0;
            Ray64 R;
            R.reset(P);
            double ct[3], nv[3];
            for (int k = 0; k < 3; k++) ct[k] = R.c[k] = o.c[k];
            R.mzz = o.mzz;
            R.ms = o.ms;
            R.calc_sit = true;
            const double noffset = (double)fminf(P.de_stop, 1.0f) * (1.0 + R.mzz * (double)P.de_stop_factor) * 0.15;
            // the centre DE gives the smooth iterations and the trap of the
            // G-buffer's colour words: without a G-buffer output, nothing
            // reads them
            float nn = 0.0f;
            if (B.gbuf) {
                R.take(de64<KIND, DEM>(P, R.c[0], R.c[1], R.c[2], R.max_it, R.calc_sit, R.ms), cnt);
                nn = R.smooth_it;
            }
            R.calc_sit = false;
            for (int q = 0; q < 3; q++) {
                const int axis = q == 0 ? 2 : q - 1;  // the order vg2, vg0, vg1
                double de[2];
                for (int side = 0; side < 2; side++) {
                    for (int k = 0; k < 3; k++) R.c[k] = ct[k] + P.vg[axis][k] * noffset;
                    if (side) {
                        for (int k = 0; k < 3; k++) R.c[k] += P.vg[axis][k] * (-2.0 * noffset);
                    }
                    if (N32) {
                        int it32 = 0;
                        float ot = 0.0f;
                        R.de = (double)de32_analytic<DEM>(P, R.c, R.ms, it32, cnt, &ot);
                        R.otrap = (double)ot;
                        cnt.nrm_32 += it32;
                    } else {
                        R.take(de64<KIND, DEM>(P, R.c[0], R.c[1], R.c[2], R.max_it, R.calc_sit, R.ms), cnt);
                    }
                    de[side] = R.de;
                }
                nv[axis] = (de[0] - de[1]) * 0.5;
            }
            for (int k = 0; k < 3; k++) R.c[k] = ct[k];
            write_surface(P, B, o.pix, R, nv, nn, (o.info & SR_DELIM) != 0, (uint16_t)(o.info & SR_ITERS), o.step_count);
            cnt.hits++;
            hit = true;
            pix = o.pix;
        }
        // the hit list: the surface records' order, -1 where a record is
        // empty (no device-wide atomic: one per wave turn serialised the
        // pass on a single counter)
        if (i < n) B.hits[i] = hit ? pix : -1;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) B.ctr[C_HITS] = (unsigned long long)n;
    flush_counters(cnt, B.ctr);
}

// create_xy_vecs_from_normals (src/math.rs:218-232): two unit vectors
// across the normal n (view frame)
__device__ inline void xy_vecs(const double n[3], double vx[3], double vy[3]) {
    const double d = n[1] * n[1] + n[0] * n[0];
    if (d < 1e-50) {
        vx[0] = 1.0, vx[1] = 0.0, vx[2] = 0.0;
        vy[0] = 0.0, vy[1] = 1.0, vy[2] = 0.0;
        return;
    }
    const double a = acos(fmin(fmax(-n[2] / sqrt(d + n[2] * n[2] + 1e-100), -1.0), 1.0)) * 0.5;
    double sin_a, cos_a;
    sincos(a, &sin_a, &cos_a);
    sin_a /= sqrt(d);
    const double dd = -n[1] * sin_a, d1 = n[0] * sin_a;
    vx[0] = 1.0 - 2.0 * d1 * d1, vx[1] = 2.0 * dd * d1, vx[2] = 2.0 * d1 * cos_a;
    vy[0] = vx[1], vy[1] = 1.0 - 2.0 * dd * dd, vy[2] = -2.0 * dd * cos_a;
}

// The surface values of the general modes at a refined surface point
// (surface_found, src/calc.rs:2471-2517): the normal -- calculate_normals
// with smooth normals 0..8 (src/calc.rs:1813-1903) or
// calculate_normals_on_smooth_it (1910-2003) -- with its roughness, NN,
// then do_color_on_it (2010-2025) and do_color (2068-2090) on the last
// evaluation's orbit.  All in f64, in the reference's order; the number of
// evaluations depends on the modes alone, so a wave's lanes stay together.
template <int KIND, int DEM>
__device__ void surface_ext(const DevP &P, const SurfRec &o, Cnt &cnt, double n[3], float &nn, float &rough,
                            uint32_t &otrap_w, int &grad5) {
    const double ct[3] = {o.c[0], o.c[1], o.c[2]};
    const float ms = o.ms;
    const int sm = P.sm_normals;
    const bool on_de = P.normals_on_de != 0;
    double noffset = (double)fminf(P.de_stop, 1.0f) * (1.0 + o.mzz * (double)P.de_stop_factor) * 0.15;
    double c[3];
    De64 last;
    // one evaluation at c: the DE, or the smooth iteration count (normals
    // on the smooth iteration); the sample's value
    auto eval = [&](int calc_sit) -> double {
        if (on_de) {
            last = de64<KIND, DEM>(P, c[0], c[1], c[2], P.max_it, calc_sit, ms);
            cnt.de64++;
        } else {
            last = mand64<KIND>(P, c[0], c[1], c[2], P.max_it);
        }
        cnt.p8_64 += last.p8;
        cnt.box_64 += last.box;
        cnt.oth_64 += last.oth;
        return on_de ? last.de : (double)last.smooth_it;
    };
    auto at = [&](const double *base) {
        c[0] = base[0];
        c[1] = base[1];
        c[2] = base[2];
    };
    auto addw = [&](const double *dir, double w) {
        for (int k = 0; k < 3; k++) c[k] += dir[k] * w;
    };
    const double sgn = on_de ? 1.0 : -1.0;  // the smooth iteration falls outwards, the DE rises
    at(ct);
    double dnn = eval(1);
    nn = last.smooth_it;
    n[0] = n[1] = n[2] = 0.0;
    if (sm == 8) {
        const double ssn = noffset * 1.3333;
        for (int a = -2; a <= 2; a++)
            for (int b = -2; b <= 2; b++)
                for (int cc = -2; cc <= 2; cc++) {
                    if ((a | b | cc) == 0) continue;
                    at(ct);
                    addw(P.vg[2], (double)a * ssn);
                    addw(P.vg[1], (double)b * ssn);
                    addw(P.vg[0], (double)cc * ssn);
                    const double d = sgn * eval(0);
                    if (a != 0) n[2] += d / (double)a;
                    if (b != 0) n[1] += d / (double)b;
                    if (cc != 0) n[0] += d / (double)cc;
                }
        for (int k = 0; k < 3; k++) n[k] *= 0.0075;
    } else {
        for (int q = 0; q < 3; q++) {
            const int axis = q == 0 ? 2 : q - 1;  // the order vg2, vg0, vg1
            at(ct);
            addw(P.vg[axis], noffset);
            const double a = eval(0);
            addw(P.vg[axis], -2.0 * noffset);
            const double b = eval(0);
            n[axis] = on_de ? (a - b) * 0.5 : (b - a) * 0.5;
        }
    }
    rough = 0.0f;
    if (sm > 0) {
        noffset *= 2.0;
        if (on_de && sm < 8) {
            at(ct);
            addw(P.vg[0], -noffset);
            dnn += eval(0);
            addw(P.vg[0], 2.0 * noffset);
            dnn += eval(0);
            at(ct);
            addw(P.vg[1], -noffset);
            dnn += eval(0);
            addw(P.vg[1], 2.0 * noffset);
            dnn = (dnn + eval(0)) * 0.2;
        } else if (!on_de) {
            double acc = (double)nn;
            const int axes[4] = {0, 0, 1, 1};
            const double sg[4] = {-1.0, 1.0, -1.0, 1.0};
            for (int q = 0; q < 4; q++) {
                at(ct);
                addw(P.vg[axes[q]], sg[q] * noffset);
                acc += eval(1);
            }
            nn = (float)(acc * 0.2);
            dnn = (double)nn;
        }
        const double ssn = noffset * 3.0 / ((double)sm + 0.5);
        const double dm = (double)sm * 2.0;
        double vx0[3], vy0[3], dirs[2][3];
        xy_vecs(n, vx0, vy0);
        for (int k = 0; k < 3; k++) {
            dirs[0][k] = vx0[0] * P.vg[0][k] + vx0[1] * P.vg[1][k] + vx0[2] * P.vg[2][k];
            dirs[1][k] = vy0[0] * P.vg[0][k] + vy0[1] * P.vg[1][k] + vy0[2] * P.vg[2][k];
        }
        double s1[2], ds[2];
        for (int w = 0; w < 2; w++) {
            at(ct);
            addw(dirs[w], -(double)sm * ssn);
            s1[w] = ds[w] = 0.0;
            for (int k = -sm; k <= sm; k++) {
                if (k != 0) {
                    const double v = eval(on_de ? 0 : 1);
                    const double d = (on_de ? v - dnn : dnn - v) / (double)k;
                    s1[w] += d;
                    ds[w] += d * d;
                }
                addw(dirs[w], ssn);
            }
        }
        const double dsg = ds[0] * dm - s1[0] * s1[0] + ds[1] * dm - s1[1] * s1[1];
        const double dt2 = noffset * 0.5 / (dm * ssn);
        const double ra = (dsg * 7.0 * dt2 * dt2 + 1e-40) / (n[0] * n[0] + n[1] * n[1] + n[2] * n[2] + 1e-40);
        rough = (float)fmin(fmax(sqrt(fmax(ra, 0.0)) - 0.05, 0.0), 1.0);  // calc_roughness (src/calc.rs:1806-1811)
        if (sm < 8) {
            n[0] += s1[0] * dt2;
            n[1] += s1[1] * dt2;
        }
    }
    // do_color_on_it: the orbit from the start point, or after
    // color_on_it - 1 iterations (with the 3D bailout)
    double v[3] = {last.v[0], last.v[1], last.v[2]};
    double rout = last.rout, rold = last.rold, otrap = last.otrap;
    if (P.color_on_it == 1) {
        v[0] = ct[0], v[1] = ct[1], v[2] = ct[2];
    } else if (P.color_on_it > 1) {
        const De64 r = de64<KIND, DEM>(P, ct[0], ct[1], ct[2], P.color_on_it - 1, 0, ms, P.rstop3d);
        cnt.de64++;
        v[0] = r.v[0], v[1] = r.v[1], v[2] = r.v[2];
        rout = r.rout, rold = r.rold, otrap = r.otrap;
    }
    // do_color
    const double pi = 3.141592653589793;
    double sc;
    grad5 = -1;
    switch (P.color_option) {
        case 1: sc = log(rout / (rold + 1.0)) * (double)P.mct_color_mul; break;
        case 2: sc = (atan2(v[1] - ct[1], v[0] - ct[0]) + pi) * 5200.0; break;
        case 3: sc = (atan2(v[2] - ct[2], v[0] - ct[0]) + pi) * 5200.0; break;
        case 4: sc = (atan2(v[2] - ct[2], v[1] - ct[1]) + pi) * 5200.0; break;
        case 5: {
            sc = (atan2(v[0], v[1]) + pi) * 5215.0;
            const double r = sqrt(v[0] * v[0] + v[1] * v[1] + 1e-100 + v[2] * v[2]);
            grad5 = (int)clip15((float)((pi + asin(fmin(fmax(v[2] / r, -1.0), 1.0)) * 2.0) * 5215.0));
            break;
        }
        default: sc = otrap * 4096.0; break;
    }
    otrap_w = clip15((float)sc);
}

// The normals pass of the general surface modes (smooth normals, normals on
// the smooth iteration, colour options 1..5, colour on an iteration):
// surface_ext per refined surface point, in f64, then the G-buffer words
// and the hit list as k_normals writes them.
template <int KIND, int DEM, bool BATCH>
__global__ __launch_bounds__(256) void k_normals_ext(DevP P0, Bufs B, const DevP *__restrict__ gtab, int nframes) {
    const LdsFrames tab = BATCH ? stage_frames(gtab, nframes) : LdsFrames{nullptr};
    Cnt cnt;
    const long long n_items = (long long)B.ctr[C_SURF];
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long base = (long long)blockIdx.x * blockDim.x + (threadIdx.x & ~63); base < n_items; base += stride) {
        const long long i = base + lane_id();
        const SurfRec o = i < n_items ? B.surf[i] : SurfRec{{0, 0, 0}, 0, 0, 0, -1, 0};
        bool hit = false;
        int pix = 0;
        if (o.pix >= 0) {
            const int fr = BATCH ? o.info >> 20 : 0;
            const DevP &P = BATCH ? tab[fr] : P0;
            double nv[3];
            float nn, rough;
            uint32_t otrap_w;
            int grad5;
            surface_ext<KIND, DEM>(P, o, cnt, nv, nn, rough, otrap_w, grad5);
            int16_t nq[3];
            store_normal(nv, nq);
            const bool de_limited = (o.info & SR_DELIM) != 0;
            float g;
            if (de_limited) {
                const float lg = (float)log((double)fmaxf(P.de_stop, o.ms) * P.step_width);
                g = 32767.0f - (nn + P.d_col_plus + P.col_var * lg) * P.mcts_m;
            } else {
                g = 32767.0f - nn * P.mcts_m;
            }
            const uint32_t si_gradient = grad5 >= 0 ? (uint32_t)grad5 : clip15(g);
            const uint32_t sh = (uint32_t)rtef(fminf(fmaxf(o.step_count, 0.0f), 1023.0f));
            Ray64 R;
            R.mzz = o.mzz;
            uint32_t zp = R.zpos_fine(P);
            if (P.sm_normals > 0) zp |= (uint32_t)rtef(rough * 255.0f) & 0xFFu;
            const double l = sqrt(nv[0] * nv[0] + nv[1] * nv[1] + nv[2] * nv[2]);
            const int px = o.pix;
            if (B.depth) B.depth[px] = (float)o.mzz;
            if (B.depth64) B.depth64[px] = o.mzz;
            if (B.iters) B.iters[px] = (uint16_t)(o.info & SR_ITERS);
            if (B.normals) {
                for (int k = 0; k < 3; k++) B.normals[3 * px + k] = l > 0 ? (float)(nv[k] / l) : 0.0f;
            }
            if (B.gbuf) {
                uint32_t *w = B.gbuf + 5 * (size_t)px;
                w[0] = (uint32_t)(uint16_t)nq[0] | ((uint32_t)(uint16_t)nq[1] << 16);
                w[1] = (uint32_t)(uint16_t)nq[2] | (sh << 16);
                w[2] = zp;
                w[3] = si_gradient | (otrap_w << 16);
                w[4] = 5000;
            }
            cnt.hits++;
            hit = true;
            pix = px;
        }
        // the hit list: the surface records' order, -1 where a record is
        // empty (no device-wide atomic: one per wave turn serialised the
        // pass on a single counter)
        if (i < n_items) B.hits[i] = hit ? pix : -1;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) B.ctr[C_HITS] = (unsigned long long)n_items;
    flush_counters(cnt, B.ctr);
}

// ---------------------------------------------------------------------------
// secondary rays: SrTracer::march_ray (src/reflect.rs:1464-1674), the march
// of a reflected ray from a surface point, as the reference's card batches
// them (gpu::reflect_batch's march pipeline, src/gpu.rs:2521-2737).  One
// task per lane, one f64 DE site per turn: the first DE, the march, the
// bisection on the ray length (or bin_search_it when the iteration limit
// stopped it), then the normal's seven DEs.

enum RayPhase : int { RP_FIRST, RP_MARCH, RP_BIS, RP_BSI, RP_NRM };
enum : int32_t { RH_HIT = 1, RH_DE_LIMITED = 2 };

// the hit's card record (the surface part of march_ray)
__device__ inline void ray_hit_record(const DevP &P, const Ray64 &R, const double n[3], float nn, bool de_limited,
                                      double zz2, float last_zz, bool depth_sqrt, float step_count,
                                      mb3d_hip_ray_hit &h) {
    int16_t nq[3];
    store_normal(n, nq);
    float g;
    if (de_limited) {
        const float lg = (float)log((double)R.ms * P.step_width);
        g = 32767.0f - (nn + P.d_col_plus + P.col_var * lg) * P.mcts_m;
    } else {
        g = 32767.0f - nn * P.mcts_m;
    }
    const uint32_t si_gradient = clip15(g);
    const uint32_t otrap = clip15((float)(R.otrap * 4096.0));  // do_color, option 0
    double ds = zz2;
    if (depth_sqrt) {
        const double a = zz2 + (double)last_zz;
        ds = sqrt(fmax(a * a - (double)last_zz + 1e-30, 0.0));
    }
    const long long v = (long long)rte(P.zc_mul * (sqrt(fmax(ds, 0.0) * P.zcorr + 1.0) - 1.0));
    const uint32_t z = (uint32_t)min(max(8388352ll - v, 0ll), 8388352ll);
    const uint32_t sh = (uint32_t)rtef(fminf(fmaxf(step_count, 0.0f), 1023.0f));
    h.gbuf[0] = (uint32_t)(uint16_t)nq[0] | ((uint32_t)(uint16_t)nq[1] << 16);
    h.gbuf[1] = (uint32_t)(uint16_t)nq[2] | (sh << 16);
    h.gbuf[2] = z << 8;
    h.gbuf[3] = si_gradient | (otrap << 16);
    h.gbuf[4] = 5000;
}

// (three waves per SIMD, 168 VGPRs: 2.7 ms against 2.9 ms at the compiler's
// two on the 1080p bulb's reflection batch; four spill, 3.5 ms)
template <int KIND, int DEM>
__global__ __launch_bounds__(256, 3) void k_rays(DevP P, const mb3d_hip_ray_task *__restrict__ tasks, long long n,
                                              int depth_sqrt, mb3d_hip_ray_hit *__restrict__ hits,
                                              unsigned long long *__restrict__ ctr) {
    Cnt cnt;
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += stride) {
        const mb3d_hip_ray_task T = tasks[t];
        mb3d_hip_ray_hit h;
        const double dir[3] = {(double)T.dir[0], (double)T.dir[1], (double)T.dir[2]};
        const int px = (int)(T.pixel % (uint32_t)P.width), py = (int)(T.pixel / (uint32_t)P.width);
        double cs[3], vf[3];
        pixel_ray(P, px, py, cs, vf);
        const double vv = vf[0] * vf[0] + vf[1] * vf[1] + vf[2] * vf[2];
        const double dd = dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2];
        const double zzplus = (vf[0] * dir[0] + vf[1] * dir[1] + vf[2] * dir[2]) / sqrt(vv * dd + 1e-100);
        double zz = (double)T.zz;
        const float last_zz = (float)fabs(zz);
        double zz2 = fmin((double)P.de_stop * 0.25, 1.0);
        zz += zz2 * zzplus;
        Ray64 R;
        R.reset(P);
        for (int k = 0; k < 3; k++) R.c[k] = T.origin[k] + dir[k] * zz2;
        auto ms_at = [&](double z) { return (float)((double)P.de_stop * (1.0 + fabs(z) * (double)P.de_stop_factor)); };
        R.ms = ms_at(zz);
        const double max_l = T.max_len > 0.0f ? (double)T.max_len : P.zend;
        bool open_air = true, de_limited = true;
        float step_count = 0.0f, rsfmul = 1.0f;
        double dtmp = 0.0, rlast_de = 0.0, rlast_sw = 0.0, zz2_paint = 0.0, bis_ds = 0.0, bsi_before = 0.0;
        double ct[3] = {0, 0, 0}, nv[3] = {0, 0, 0}, de_a = 0.0, noffset = 0.0;
        float nn = 0.0f;
        int bis_left = 0, nq = 0;
        uint32_t guard = 0;
        Refine F;
        int phase = RP_FIRST;
        if (T.inside) {
            zz2 = P.zend;
        } else {
            open_air = false;
            for (;;) {
                R.take(de64<KIND, DEM>(P, R.c[0], R.c[1], R.c[2], R.max_it, R.calc_sit, R.ms), cnt);
                bool refine = false;
                if (phase == RP_FIRST) {
                    dtmp = R.de;
                    rlast_sw = dtmp * (double)P.z_step_div;
                    if (!(dtmp > 1e-10)) refine = true;
                } else if (phase == RP_MARCH) {
                    dtmp = R.de;
                    if (R.it >= R.max_its) {
                        de_limited = false;
                        refine = true;
                    } else if (dtmp < (double)R.ms && R.it >= P.min_it) {
                        refine = true;
                    } else {
                        if (dtmp > rlast_de + rlast_sw) dtmp = rlast_de + rlast_sw;
                        if (rlast_de > dtmp + 1e-10) {
                            const double d = rlast_sw / (rlast_de - dtmp);
                            rsfmul = d < 1.0 ? (float)fmax(d, 0.5) : 1.0f;
                        } else {
                            rsfmul = 1.0f;
                        }
                        if (++guard > 1000000u) {
                            open_air = true;
                            break;
                        }
                    }
                } else if (phase == RP_BIS) {
                    bis_ds = R.de < (double)R.ms ? fabs(bis_ds) * -0.55 : fabs(bis_ds) * 0.55;
                    goto bis_head;
                } else if (phase == RP_BSI) {
                    if (F.after(R, P)) continue;
                    zz2 = R.mzz;
                    zz += (zz2 - bsi_before) * zzplus;
                    goto normals;
                } else {  // RP_NRM: the centre (smooth iterations), then vg2, vg0, vg1 both ways
                    if (nq == 0) {
                        nn = R.smooth_it;
                        h.iters = R.it;
                        R.calc_sit = false;
                    } else {
                        const int axis = ((nq - 1) >> 1) == 0 ? 2 : ((nq - 1) >> 1) - 1;
                        if ((nq & 1) == 1) {
                            de_a = R.de;
                        } else {
                            nv[axis] = (de_a - R.de) * 0.5;
                        }
                    }
                    if (++nq == 7) {
                        for (int k = 0; k < 3; k++) R.c[k] = ct[k];
                        ray_hit_record(P, R, nv, nn, de_limited, zz2, last_zz, depth_sqrt != 0, step_count, h);
                        cnt.hits++;
                        break;
                    }
                    {
                        // the point off the centre, then the other side from it (add_weight
                        // by -2 noffset, as the reference steps)
                        const int axis = ((nq - 1) >> 1) == 0 ? 2 : ((nq - 1) >> 1) - 1;
                        if (nq & 1) {
                            for (int k = 0; k < 3; k++) R.c[k] = ct[k] + P.vg[axis][k] * noffset;
                        } else {
                            for (int k = 0; k < 3; k++) R.c[k] += P.vg[axis][k] * (-2.0 * noffset);
                        }
                    }
                    continue;
                }
                if (refine) {
                    zz2_paint = zz2;
                    if (P.de_add_steps != 0) {
                        if (de_limited) {
                            bis_left = P.de_add_steps;
                            bis_ds = rlast_sw * -0.5;
                            goto bis_head;
                        }
                        bsi_before = zz2;
                        for (int k = 0; k < 3; k++) R.v[k] = dir[k];
                        R.mzz = zz2;
                        F.start_bsi(R, P.de_add_steps);
                        phase = RP_BSI;
                        continue;
                    }
                    goto normals;
                }
                {
                    // one step of the march (src/reflect.rs:1486-1530)
                    rlast_de = dtmp;
                    double dstep = fmax((dtmp - (double)P.ms_de_sub * (double)R.ms) * (double)P.z_step_div * (double)rsfmul,
                                        0.11);
                    const double s = (double)(fmaxf(R.ms, 0.4f) * P.mh04zsd);
                    const bool count = P.dfog_on_it == 0 || R.it == P.dfog_on_it;
                    if (s < dstep) {
                        if (count) step_count += (float)(s / dstep);
                        dstep = s;
                    } else if (count) {
                        step_count += 1.0f;
                    }
                    rlast_sw = dstep;
                    zz2 += dstep;
                    cnt.steps++;
                    if (zz2 > max_l) {
                        open_air = true;
                        break;
                    }
                    for (int k = 0; k < 3; k++) R.c[k] += dir[k] * dstep;
                    zz += dstep * zzplus;
                    R.ms = ms_at(zz);
                    phase = RP_MARCH;
                    continue;
                }
            bis_head:
                // the bisection on the ray length (src/reflect.rs:1555-1571)
                if (fabs(R.de - (double)R.ms) > 0.001) {
                    zz2 += bis_ds;
                    zz += bis_ds * zzplus;
                    for (int k = 0; k < 3; k++) R.c[k] += dir[k] * bis_ds;
                    R.ms = ms_at(zz);
                    if (--bis_left > 0) {
                        phase = RP_BIS;
                        continue;
                    }
                }
            normals:
                // calculate_normals at the point (src/calc.rs:1813-1858)
                R.mzz = fabs(zz);
                noffset = (double)fminf(P.de_stop, 1.0f) * (1.0 + R.mzz * (double)P.de_stop_factor) * 0.15;
                for (int k = 0; k < 3; k++) ct[k] = R.c[k];
                R.calc_sit = true;
                nq = 0;
                phase = RP_NRM;
            }
        }
        if (open_air) {
            const uint32_t sh = (uint32_t)rtef(fminf(fmaxf(step_count, 0.0f), 1023.0f));
            h.gbuf[0] = 0;
            h.gbuf[1] = sh << 16;
            h.gbuf[2] = 32768u << 16;
            h.gbuf[3] = 0;
            h.gbuf[4] = 0;
            zz2_paint = zz2;
        }
        h.flags = (open_air ? 0 : RH_HIT) | (de_limited ? RH_DE_LIMITED : 0);
        for (int k = 0; k < 3; k++) h.pos[k] = R.c[k];
        h.zz = zz;
        h.zz2 = zz2;
        h.zz2_paint = zz2_paint;
        h.step_count = step_count;
        if (open_air) h.iters = R.it;
        hits[t] = h;
    }
    flush_counters(cnt, ctr);
}

// The shadow and occlusion pass at the rays' hits (calc_hs and deao_1pos,
// src/reflect.rs:805-1090; the reference's card runs it as a pipeline of its
// own after the march, src/gpu.rs:2562-2575).  The AO ray sets of the four
// qualities and the pairwise overlap term acos(d_k . d_k2) abr are built on
// the host: they depend on the directions only.
struct RayPost {
    int n_lights, soft, ao, reserved;
    int bit[MB3D_HIP_MAX_LIGHTS], positional[MB3D_HIP_MAX_LIGHTS], visible[MB3D_HIP_MAX_LIGHTS];
    float lmax[MB3D_HIP_MAX_LIGHTS];
    double vec[MB3D_HIP_MAX_LIGHTS][3];
    float soft_radius, hs_mul, ao_max_len, pad;
    int count[4], first[4], pen_off[4];
    float dirs[26][3];
    double d_step_mul[4], de_mul[4];
    float corr_w[4];
    float pen[236];  // per quality, count x count: acos(d_k . d_k2) abr, +inf where the pair is not near enough
};
constexpr int kRayAoMax = 13;

__device__ inline void normal_matrix(const float n[3], float mo[3][3]);  // (with the DE AO pass below)

enum RayPostPhase : int { RQ_HS_FIRST, RQ_HS, RQ_AO };

// (three waves per SIMD, 168 VGPRs: about 19 ms against 22 ms at one wave
// on the 1080p bulb's reflection hits with one hard shadow and quality-2 AO)
template <int KIND, int DEM>
__global__ __launch_bounds__(256, 3) void k_ray_post(DevP P, const mb3d_hip_ray_task *__restrict__ tasks, long long n,
                                                  RayPost J, mb3d_hip_ray_hit *__restrict__ hits,
                                                  unsigned long long *__restrict__ ctr) {
    __shared__ float s_min_ra[kRayAoMax][256];
    float *min_ra = &s_min_ra[0][threadIdx.x];  // min_ra[k * 256]
    Cnt cnt;
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += stride) {
        mb3d_hip_ray_hit h = hits[t];
        if (!(h.flags & RH_HIT)) continue;
        const mb3d_hip_ray_task T = tasks[t];
        const int px = (int)(T.pixel % (uint32_t)P.width), py = (int)(T.pixel / (uint32_t)P.width);
        double cs[3], vf[3];
        pixel_ray(P, px, py, cs, vf);
        const double sw = P.step_width;
        const double ic[3] = {h.pos[0], h.pos[1], h.pos[2]};
        const double mzz = fabs(h.zz);
        const float zzf = (float)mzz;
        const int16_t nq[3] = {(int16_t)(h.gbuf[0] & 0xFFFFu), (int16_t)(h.gbuf[0] >> 16), (int16_t)(h.gbuf[1] & 0xFFFFu)};
        uint32_t shadow = h.gbuf[1] >> 16;
        uint32_t amb = h.gbuf[4] & 0xFFFFu;
        Ray64 R;
        R.reset(P);
        for (int k = 0; k < 3; k++) R.c[k] = ic[k];
        auto ms_at = [&](double z) { return (float)((double)P.de_stop * (1.0 + fabs(z) * (double)P.de_stop_factor)); };
        // calc_hs's values
        const double max_lhs = (double)((float)(P.width + P.height) * 0.6f *
                                        (1.0f + 0.5f * fminf(zzf, (float)P.zend * 0.4f) * fmaxf((float)P.fov_y, 0.0f) /
                                                    (float)P.height) *
                                        J.hs_mul);
        double nvec[3];
        {
            const double nn[3] = {(double)nq[0] / 32767.0, (double)nq[1] / 32767.0, (double)nq[2] / 32767.0};
            for (int k = 0; k < 3; k++) nvec[k] = nn[0] * P.vg[0][k] + nn[1] * P.vg[1][k] + nn[2] * P.vg[2][k];
        }
        int li = J.soft ? J.n_lights - 1 : 0;
        const int li_end = J.n_lights;
        double hsvec[3] = {0, 0, 0}, dt1 = 0.0, dmaxl = 0.0, zz2 = 0.0, zz2mul = 0.0, d = 0.0, last_de = 0.0, st = 0.0;
        float rsfd = 1.0f, zrsoft = 1.0f, zrs_mul = 0.0f;
        int guard = 0;
        int phase = -1;
        if (J.n_lights > 0 && !J.soft) {
            uint32_t mask = 0;
            for (int i = 0; i < J.n_lights; i++) mask |= 0x400u << J.bit[i];
            shadow &= ~mask;
        }
        // AO state
        int q = 0, nd = 0, k_ao = 0, gao = 0;
        double ms_ao = 0.0, step_ao = 0.0, max_dist = 0.0, md_d10 = 0.0, adt = 0.0, as = 1.0, sv[3] = {0, 0, 0};
        bool aend = false;
        float rw[3][3];
        for (;;) {
            if (phase < 0) {
                // the next light's march, or the occlusion when the lights are done
                bool started = false;
                while (!started && li < li_end && J.n_lights > 0) {
                    const int i = li++;
                    if (J.soft) shadow |= 0xFC00u;  // lit, unless the march says otherwise
                    for (int k = 0; k < 3; k++) R.c[k] = ic[k];
                    zz2 = mzz;
                    R.ms = ms_at((double)zzf);
                    double len = max_lhs;
                    for (int k = 0; k < 3; k++) hsvec[k] = J.vec[i][k];
                    bool reach = true;
                    if (J.positional[i]) {
                        double v[3];
                        for (int k = 0; k < 3; k++) v[k] = J.vec[i][k] - R.c[k];
                        const double dd2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
                        if (dd2 > (double)(J.lmax[i] * J.hs_mul)) {
                            reach = false;
                        } else {
                            if (dd2 < (len * sw) * (len * sw)) len = sqrt(dd2) / sw;
                            const double l = sqrt(dd2);
                            for (int k = 0; k < 3; k++) hsvec[k] = v[k] * (-sw / l);
                        }
                    }
                    const uint32_t bit = 0x400u << J.bit[i];
                    if (!reach) {
                        if (!J.soft) shadow |= bit;
                        continue;
                    }
                    zz2mul = -(hsvec[0] * vf[0] + hsvec[1] * vf[1] + hsvec[2] * vf[2]) /
                             sqrt((hsvec[0] * hsvec[0] + hsvec[1] * hsvec[1] + hsvec[2] * hsvec[2]) *
                                  (vf[0] * vf[0] + vf[1] * vf[1] + vf[2] * vf[2]));
                    if (J.soft) {
                        const float sr = fmaxf(J.soft_radius, 0.001f);
                        zrs_mul = J.positional[i] ? ((J.visible[i] & 6) == 2 ? 70.0f / sr : 40.0f / sr) : 80.0f / sr;
                    }
                    if (!(len > 0.0)) continue;
                    if (nvec[0] * hsvec[0] + nvec[1] * hsvec[1] + nvec[2] * hsvec[2] > 0.0) {
                        if (J.soft)
                            shadow &= 0x3FFu;
                        else
                            shadow |= bit;
                        continue;
                    }
                    dt1 = dmaxl = len;
                    rsfd = 1.0f;
                    zrsoft = 1.0f;
                    guard = 0;
                    phase = RQ_HS_FIRST;
                    started = true;
                }
                if (!started) {
                    for (int k = 0; k < 3; k++) R.c[k] = ic[k];
                    if (!J.ao) break;
                    // deao_1pos's set-up
                    q = min(max(T.ao_q, 0), 3);
                    nd = J.count[q];
                    R.calc_sit = false;
                    ms_ao = (double)fminf((float)((double)P.de_stop * (1.0 + mzz * (double)P.de_stop_factor)), 1e6f);
                    const double s_max_d = (double)J.ao_max_len * 0.5 *
                                           sqrt((double)P.height * (double)P.height + (double)P.width * (double)P.width);
                    step_ao = ms_ao / (double)P.de_stop;
                    max_dist = s_max_d * sqrt(step_ao);
                    float nf[3] = {(float)nq[0], (float)nq[1], (float)nq[2]};
                    const float l = fmaxf(sqrtf(nf[0] * nf[0] + nf[1] * nf[1] + nf[2] * nf[2]), 1e-30f);
                    for (int k = 0; k < 3; k++) nf[k] = nf[k] / l;
                    normal_matrix(nf, rw);
                    const double dsm = J.d_step_mul[q];
                    if (P.de_stop_factor != 0.0f)
                        ms_ao /= dsm * dsm;
                    else
                        ms_ao = (double)P.de_stop / (dsm * dsm);
                    md_d10 = 0.1 / (max_dist * J.de_mul[q]);
                    k_ao = 0;
                    phase = RQ_AO;
                    goto ao_ray;
                }
            }
            if (phase == RQ_AO && k_ao >= nd) break;
            R.take(de64<KIND, DEM>(P, R.c[0], R.c[1], R.c[2], R.max_it, R.calc_sit, R.ms), cnt);
            if (phase == RQ_HS_FIRST) {
                d = R.de;
                phase = RQ_HS;
                goto hs_step;
            } else if (phase == RQ_HS) {
                d = R.de;
                if (J.soft) {
                    const float r = (float)((dmaxl - dt1) / max_lhs);
                    const float r8 = (r * r) * (r * r);
                    zrsoft = fminf(zrsoft, (float)((d - (double)R.ms) * (double)zrs_mul / (dmaxl - dt1 + 0.11)) + r8 * r8);
                }
                if (R.it >= R.max_its || d < (double)R.ms) goto hs_done;
                if (d > last_de + st) d = last_de + st;
                if (last_de > d + 1e-30) {
                    const float tt = (float)(st / (last_de - d));
                    rsfd = tt < 1.0f ? fmaxf(tt, 0.5f) : 1.0f;
                } else {
                    rsfd = 1.0f;
                }
                if (dt1 < 0.0 || guard > 1000000) goto hs_done;
                goto hs_step;
            } else {
                // an AO probe's DE (deao_1pos's ray loop)
                const double d2 = R.de;
                as = fmin(as, (d2 - ms_ao + adt * md_d10) / adt);
                if (as < 0.02) goto ao_next;
                adt += fmax(d2, adt * J.d_step_mul[q]);
                if (aend || gao > 100000) goto ao_next;
                goto ao_probe;
            }
        hs_step:
            guard += 1;
            last_de = d;
            st = (double)fminf((float)(d * (double)P.z_step_div * (double)rsfd), fmaxf(R.ms, 0.4f) * P.mh04zsd);
            dt1 -= st;
            for (int k = 0; k < 3; k++) R.c[k] += hsvec[k] * (-st);
            zz2 += st * zz2mul;
            R.ms = ms_at(zz2);
            cnt.steps++;
            continue;
        hs_done:
            if (J.soft) {
                shadow = (shadow & 0x3FFu) | ((uint32_t)rtef(fminf(fmaxf(zrsoft, 0.0f), 1.0f) * 63.4f) << 10);
                li = li_end;
            } else if (dt1 > 0.0) {
                shadow |= 0x400u << J.bit[li - 1];
            }
            phase = -1;
            continue;
        ao_next:
            min_ra[k_ao * 256] = (float)(fmax(as, 0.0) * J.de_mul[q]);
            k_ao++;
        ao_ray:
            if (k_ao >= nd) continue;
            {
                const float *dr = J.dirs[J.first[q] + k_ao];
                const float v[3] = {rw[0][0] * dr[0] + rw[0][1] * dr[1] + rw[0][2] * dr[2],
                                    rw[1][0] * dr[0] + rw[1][1] * dr[1] + rw[1][2] * dr[2],
                                    rw[2][0] * dr[0] + rw[2][1] * dr[1] + rw[2][2] * dr[2]};
                // rotate_vector_reverse by vgrads rows normalised to the step width
                for (int k = 0; k < 3; k++) sv[k] = 0.0;
                for (int r = 0; r < 3; r++) {
                    const double vv = sqrt(P.vg[r][0] * P.vg[r][0] + P.vg[r][1] * P.vg[r][1] + P.vg[r][2] * P.vg[r][2]);
                    const double sc = sw / vv;
                    for (int k = 0; k < 3; k++) sv[k] += (double)v[r] * (P.vg[r][k] * sc);
                }
                adt = step_ao * J.d_step_mul[q];
                as = 1.0;
                aend = false;
                gao = 0;
            }
        ao_probe:
            gao += 1;
            if (adt > max_dist) {
                adt = max_dist;
                aend = true;
            }
            for (int k = 0; k < 3; k++) R.c[k] = ic[k] + sv[k] * adt;
        }
        if (J.ao) {
            // the rays' overlap correction and the AO word
            double amount = 0.0;
            const float *pen = J.pen + J.pen_off[q];
            for (int k = 0; k < nd; k++) {
                const float mk = min_ra[k * 256];
                float s_add = 0.0f;
                if (mk < 1.0f) {
                    const float max_add = 1.0f - mk;
                    for (int k2 = 0; k2 < nd; k2++) {
                        if (k2 == k) continue;
                        const float overlap = min_ra[k2 * 256] - pen[k * nd + k2] + 1.0f;
                        if (overlap > 0.0f) s_add += fminf(max_add, overlap) * J.corr_w[q];
                    }
                }
                amount += (double)fminf(s_add + mk, 1.0f);
            }
            amb = (uint32_t)fmax(rte(16383.0 * (1.0 - amount / (double)nd)), 0.0);
        }
        h.gbuf[1] = (h.gbuf[1] & 0xFFFFu) | ((shadow & 0xFFFFu) << 16);
        h.gbuf[4] = amb;
        hits[t] = h;
    }
    flush_counters(cnt, ctr);
}

// ---------------------------------------------------------------------------
// the post passes: hard shadows and DE AO, same state-machine shape

// The DE of a post ray: the bulk precision, and f64 where the decision is
// close (within the guard band of the DE stop, at the iteration limit, or an
// FP32 overflow), so that no DE site is duplicated per state.
template <int KIND, int DEM, bool F32>
__device__ __forceinline__ void post_de(const DevP &P, Ray64 &R, bool want64, Cnt &cnt) {
    if constexpr (F32) {
        if (!want64) {
            int it = 0;
            const float d32 = de32<KIND, DEM>(P, R.c, R.ms, it, cnt);
            if (d32 < __builtin_inff() && it < R.max_it && d32 > R.ms * P.guard) {
                R.de = (double)d32;
                R.it = it;
                R.max_its = R.max_it;
                return;
            }
            if (!(d32 < __builtin_inff())) cnt.overflow++;
        }
    }
    R.take(de64<KIND, DEM>(P, R.c[0], R.c[1], R.c[2], R.max_it, R.calc_sit, R.ms), cnt);
}

// hs_surface_start / surface_point: the start of the f64 surface point of a
// G-buffer pixel from its depth code (src/calc.rs:2741-2753, 2163-2171)
__device__ __forceinline__ void ray_from_code(const DevP &P, Ray64 &R, int x, int y, uint32_t zpos_fine) {
    R.reset(P);
    pixel_ray(P, x, y, R.c, R.v);
    const double zf = (double)(zpos_fine >> 8);
    R.mzz = depth_of_code(P, zf, 8388351.5);
    for (int k = 0; k < 3; k++) R.c[k] += R.v[k] * R.mzz;
    R.upd(P);
}

enum ShPhase : int { SH_SURF, SH_REF, SH_D0, SH_DS };

template <int KIND, int DEM, bool F32>
// (three waves per SIMD: the post pass of the 1080p bulb with one light and
// quality-1 AO took 7.7 ms against 8.7 ms at the compiler's two, identical
// records; with k_deao below)
__global__ __launch_bounds__(256, 3) void k_shadow(DevP P, Bufs B, ShadowJob J) {
    Cnt cnt;
    const long long n_items = (long long)B.ctr[C_HITS];
    uint32_t clear_mask = 0;
    for (int l = 0; l < J.n; l++) clear_mask |= (uint32_t)J.bit[l];
    bool have = false, exhausted = false;
    Ray64 R;
    Refine F;
    int pix = 0, phase = SH_SURF, l = 0, guard = 0, y = 0;
    uint32_t sh = 0, zcode = 0;
    double nvec[3] = {0, 0, 0}, ic[3] = {0, 0, 0};
    double max_lhs = 0, zz2 = 0, zz2mul = 0, dt1 = 0, d = 0, last_de = 0, st = 0;
    float rsf = 2.0f;
    WaveChunk pq;
    chunk_init(pq);
    for (;;) {
        bool need = !have && !exhausted;
        for (int tries = 0; tries < 4 && __any(need); tries++) {
            const long long slot = wave_take(need, &B.ctr[C_POSTQ], pq, 64, n_items);
            if (!need) continue;
            if (slot < 0) {
                exhausted = true;
                need = false;
                continue;
            }
            pix = B.hits[slot];
            if (pix < 0) continue;
            need = false;
            const uint32_t *w = B.gbuf + 5 * (size_t)pix;
            if ((w[2] >> 16) < 32768u && (w[3] & 0xFFFF) < 32768u) {
                have = true;
                sh = (w[1] >> 16) & ~clear_mask;
                const double n0 = (double)(int16_t)(w[0] & 0xFFFF) / 32767.0;
                const double n1 = (double)(int16_t)(w[0] >> 16) / 32767.0;
                const double n2 = (double)(int16_t)(w[1] & 0xFFFF) / 32767.0;
                for (int k = 0; k < 3; k++) nvec[k] = n0 * P.vg[0][k] + n1 * P.vg[1][k] + n2 * P.vg[2][k];
                const int x = pix % P.width;
                y = P.row0 + pix / P.width;
                zcode = w[2];
                ray_from_code(P, R, x, y, zcode);
                phase = SH_SURF;
            }
        }
        if (__all(exhausted)) break;
        if (!have) continue;
        post_de<KIND, DEM, F32>(P, R, phase == SH_SURF || phase == SH_REF, cnt);
        bool done = false;
        switch (phase) {
            case SH_SURF: {
                const bool de_limited = R.it < R.max_its || R.de < (double)R.ms;
                const double zf = (double)(zcode >> 8);
                const bool more = de_limited ? F.start_bs(R, P, R.de, depth_of_code(P, zf, 8388351.9) - R.mzz,
                                                          P.de_add_steps)
                                             : F.start_bsi(R, P.de_add_steps);
                if (more) {
                    phase = SH_REF;
                    goto end;
                }
                goto start;
            }
            case SH_REF:
                if (F.after(R, P)) goto end;
                goto start;
            case SH_D0:
                d = R.de;
                goto head;
            default:  // SH_DS
                d = R.de;
                if (R.it >= R.max_its || d <= (double)R.ms) goto endl;
                if (d > last_de + st) d = last_de + st;
                if (last_de > d + 1e-30) {
                    const float t = (float)(st / (last_de - d));
                    rsf = t < 1.0f ? fmaxf(t, 0.5f) : 1.0f;
                } else {
                    rsf = 1.0f;
                }
                if (dt1 < 0.0) goto endl;
                goto head;
        }
    start: {
        // hard_shadow_row from the surface point (src/calc.rs:2793-2800)
        R.mzz = fmax(R.mzz - 0.1, 0.0);
        for (int k = 0; k < 3; k++) R.c[k] += R.v[k] * -0.1;
        R.upd(P);
        for (int k = 0; k < 3; k++) ic[k] = R.c[k];
        // hs_max_len (src/calc.rs:2767-2775)
        const float mz = (float)fmin(R.mzz, P.zend * 0.4);
        max_lhs = (double)((float)(P.width + y) * 0.6f *
                           (1.0f + 0.5f * mz * (float)fmax(P.fov_y, 0.0) / (float)P.height) * J.max_len_mul);
        l = 0;
    }
    light:
        if (l >= J.n) goto write;
        {
            const double *lv = J.vec[l];
            for (int k = 0; k < 3; k++) R.c[k] = ic[k];
            zz2 = R.mzz;
            const double ll = lv[0] * lv[0] + lv[1] * lv[1] + lv[2] * lv[2];
            const double vv = R.v[0] * R.v[0] + R.v[1] * R.v[1] + R.v[2] * R.v[2];
            zz2mul = -((lv[0] * R.v[0] + lv[1] * R.v[1] + lv[2] * R.v[2]) / fmax(sqrt(ll * vv), 1e-300));
            if (nvec[0] * lv[0] + nvec[1] * lv[1] + nvec[2] * lv[2] > 0.0) {
                sh |= (uint32_t)J.bit[l];
                l++;
                goto light;
            }
            dt1 = max_lhs;
            if (!(dt1 > 0.0)) goto endl;
            R.ms = (float)((double)P.de_stop * (1.0 + fabs(zz2) * (double)P.de_stop_factor));
            rsf = 2.0f;
            guard = 0;
            phase = SH_D0;
            goto end;
        }
    head:
        if (++guard > 1000000) goto endl;
        {
            const double *lv = J.vec[l];
            last_de = d;
            const float a = fmaxf(0.11f, (float)((d - (double)P.ms_de_sub * (double)R.ms) * (double)P.z_step_div *
                                                 (double)rsf));
            const float b = fmaxf(R.ms, 0.4f) * P.mh04zsd;
            st = (double)fminf(a, b);
            dt1 -= st;
            for (int k = 0; k < 3; k++) R.c[k] += lv[k] * -st;
            zz2 += st * zz2mul;
            R.ms = (float)((double)P.de_stop * (1.0 + fabs(zz2) * (double)P.de_stop_factor));
            cnt.steps++;
            phase = SH_DS;
            goto end;
        }
    endl:
        if (dt1 > 0.0) sh |= (uint32_t)J.bit[l];
        l++;
        goto light;
    write: {
        uint32_t *w = B.gbuf + 5 * (size_t)pix;
        w[1] = (w[1] & 0xFFFFu) | (sh << 16);
        done = true;
    }
    end:
        if (done) have = false;
    }
    flush_counters(cnt, B.ctr);
}

// normal_matrix (src/deao.rs:42-58)
__device__ inline void normal_matrix(const float n[3], float mo[3][3]) {
    const double a = acos(fmin(fmax(-(double)n[2], -1.0), 1.0)) * 0.5;
    double sa = sin(a);
    const double ca = cos(a);
    const double nn = sqrt((double)(n[1] * n[1] + n[0] * n[0]));
    double q0 = 0, q1 = 0, q2 = 0;
    if (!(nn < 1e-25)) {
        sa /= nn;
        q0 = -(double)n[1] * sa;
        q1 = (double)n[0] * sa;
    }
    const double q3 = ca;
    mo[0][0] = (float)(1.0 - 2.0 * (q1 * q1 + q2 * q2));
    mo[0][1] = (float)(2.0 * (q0 * q1 + q2 * q3));
    mo[0][2] = (float)(2.0 * (q0 * q2 - q1 * q3));
    mo[1][0] = (float)(2.0 * (q0 * q1 - q2 * q3));
    mo[1][1] = (float)(1.0 - 2.0 * (q0 * q0 + q2 * q2));
    mo[1][2] = (float)(2.0 * (q2 * q1 + q0 * q3));
    mo[2][0] = (float)(2.0 * (q0 * q2 + q1 * q3));
    mo[2][1] = (float)(2.0 * (q1 * q2 - q0 * q3));
    mo[2][2] = (float)(1.0 - 2.0 * (q0 * q0 + q1 * q1));
}

enum AoPhase : int { AO_SURF, AO_REF, AO_D };

template <int KIND, int DEM, bool F32>
// (three waves per SIMD: with quality-2 AO the post pass took 18.9 ms
// against 22.1 ms at the compiler's two)
__global__ __launch_bounds__(256, 3) void k_deao(DevP P, Bufs B, AoRays A) {
    __shared__ float s_min_ra[256 * 33];
    __shared__ AoSet s_sets[9];
    float *min_ra = s_min_ra + threadIdx.x;  // min_ra[k * 256]
    {
        const int nsets = A.side * A.side;
        const float *src = (const float *)A.sets;
        float *dst = (float *)s_sets;
        for (int i = threadIdx.x; i < nsets * (int)(sizeof(AoSet) / 4); i += blockDim.x) dst[i] = src[i];
        __syncthreads();
    }
    Cnt cnt;
    const long long n_items = (long long)B.ctr[C_HITS];
    const double sw = P.step_width;
    const double s_max_d =
        (double)A.max_len * 0.5 * sqrt((double)P.height * (double)P.height + (double)P.width * (double)P.width);
    const double de_stop = (double)P.de_stop;
    const double dsm = (double)A.d_step_mul;
    // the pixel's ray set (its dither cell) and what follows from its count
    const AoSet *set = s_sets;
    int rc = 0;
    double de_mul = 0.0, abr = 0.0;
    uint32_t rs = 0;  // the row generator's state (the random first step)
    // the camera rows scaled to the step width (normalise_matrix_to)
    double rot_v[3][3];
    for (int i = 0; i < 3; i++) {
        const double dd = sw / sqrt(P.vg[i][0] * P.vg[i][0] + P.vg[i][1] * P.vg[i][1] + P.vg[i][2] * P.vg[i][2]);
        for (int j = 0; j < 3; j++) rot_v[i][j] = P.vg[i][j] * dd;
    }
    bool have = false, exhausted = false;
    Ray64 R;
    Refine F;
    int pix = 0, phase = AO_SURF, k = 0, guard = 0;
    uint32_t zcode = 0;
    float rw[3][3];
    double ic[3] = {0, 0, 0}, sv[3] = {0, 0, 0};
    double ms_ao = 0, step_ao = 0, max_dist = 0, md_d10 = 0, dt1 = 0, s_tmp = 0;
    bool b_end = false;
    WaveChunk aq;
    chunk_init(aq);
    for (;;) {
        bool need = !have && !exhausted;
        for (int tries = 0; tries < 4 && __any(need); tries++) {
            const long long slot = wave_take(need, &B.ctr[C_AOQ], aq, 64, n_items);
            if (!need) continue;
            if (slot < 0) {
                exhausted = true;
                need = false;
                continue;
            }
            pix = B.hits[slot];
            if (pix < 0) continue;
            need = false;
            uint32_t *w = B.gbuf + 5 * (size_t)pix;
            if ((w[2] >> 16) < 32768u && (w[3] & 0xFFFF) < 32768u) {
                have = true;
                zcode = w[2];
                const float nf[3] = {(float)(int16_t)(w[0] & 0xFFFF), (float)(int16_t)(w[0] >> 16),
                                     (float)(int16_t)(w[1] & 0xFFFF)};
                const float nl = fmaxf(sqrtf(nf[0] * nf[0] + nf[1] * nf[1] + nf[2] * nf[2]), 1e-30f);
                const float nu[3] = {nf[0] / nl, nf[1] / nl, nf[2] / nl};
                normal_matrix(nu, rw);
                const int x = pix % P.width, y = P.row0 + pix / P.width;
                ray_from_code(P, R, x, y, zcode);
                set = s_sets + (A.side > 1 ? (y % A.side) * A.side + x % A.side : 0);
                rc = set->count;
                de_mul = sqrt((double)rc * 0.5);
                abr = 1.2 / asin(fmin(fmax(1.0 / de_mul, -1.0), 1.0));
                if (A.rnd) rs = A.rnd[pix];
                phase = AO_SURF;
            } else {
                w[4] = w[4] & ~0xFFFFu;
            }
        }
        if (__all(exhausted)) break;
        if (!have) continue;
        post_de<KIND, DEM, F32>(P, R, phase != AO_D, cnt);
        bool done = false;
        switch (phase) {
            case AO_SURF: {
                // surface_point (src/calc.rs:2172-2188): 5 steps to |DE - stop| <= 0.004
                const double dd = fmin(R.de, (double)R.ms * 2.0);
                const double zf = (double)(zcode >> 8);
                bool more;
                if (R.it < R.max_its || dd < (double)R.ms)
                    more = F.start_sp(R, P, dd, depth_of_code(P, zf, 8388351.9) - R.mzz, 5, 0.004);
                else
                    more = F.start_bsi(R, 5);
                if (more) {
                    phase = AO_REF;
                    goto end;
                }
                goto refined;
            }
            case AO_REF:
                if (F.after(R, P)) goto end;
                goto refined;
            default: {  // AO_D
                const double dt2 = R.de;
                s_tmp = fmin(s_tmp, (dt2 - ms_ao + dt1 * md_d10) / dt1);
                if (s_tmp < 0.02) goto ray_end;
                dt1 += fmax(dt2, dt1 * dsm);
                if (b_end || ++guard >= 100000) goto ray_end;
                goto ray_step;
            }
        }
    refined:
        if (F.kind == RK_BSI)
            R.ms = (float)((double)P.de_stop * (1.0 + fabs(R.mzz) * (double)P.de_stop_factor));
        // deao_row after surface_point (src/deao.rs:303-345)
        ms_ao = fmin((double)R.ms, 1e6);
        for (int q = 0; q < 3; q++) ic[q] = R.c[q];
        step_ao = ms_ao / de_stop;
        max_dist = s_max_d * sqrt(step_ao);
        ms_ao = P.de_stop_factor == 0.0f ? de_stop / (dsm * dsm) : ms_ao / (dsm * dsm);
        md_d10 = 0.1 / (max_dist * de_mul);
        k = 0;
    ray:
        if (k >= rc) goto write;
        {
            const float *dr = set->dir[k];
            const float v0 = rw[0][0] * dr[0] + rw[0][1] * dr[1] + rw[0][2] * dr[2];
            const float v1 = rw[1][0] * dr[0] + rw[1][1] * dr[1] + rw[1][2] * dr[2];
            const float v2 = rw[2][0] * dr[0] + rw[2][1] * dr[1] + rw[2][2] * dr[2];
            for (int q = 0; q < 3; q++)
                sv[q] = (double)v0 * rot_v[0][q] + (double)v1 * rot_v[1][q] + (double)v2 * rot_v[2][q];
            dt1 = step_ao * dsm;
            s_tmp = 1.0;
            b_end = false;
            guard = 0;
            if (A.rnd) {
                // get_rand (src/deao.rs:262-265); the first step is not held to max_dist
                rs = rs * 0x343FDu + 0x269EC3u;
                dt1 *= (double)((rs >> 8) & 0x7FFFFFu) / (double)0x7FFFFF * 1.5 + 0.5;
                goto ray_place;
            }
        }
    ray_step:
        if (dt1 > max_dist) {
            dt1 = max_dist;
            b_end = true;
        }
    ray_place:
        for (int q = 0; q < 3; q++) R.c[q] = ic[q] + sv[q] * dt1;
        cnt.steps++;
        phase = AO_D;
        goto end;
    ray_end:
        min_ra[k * 256] = (float)(fmax(s_tmp, 0.0) * de_mul);
        k++;
        goto ray;
    write: {
        double amount = 0.0;
        for (int a = 0; a < rc; a++) {
            float s_add = 0.0f;
            const float ma = min_ra[a * 256];
            if (ma < 1.0f) {
                const float max_add = 1.0f - ma;
                for (int b = 0; b < rc; b++) {
                    if (b == a) continue;
                    const float dt = set->dir[a][0] * set->dir[b][0] + set->dir[a][1] * set->dir[b][1] +
                                     set->dir[a][2] * set->dir[b][2];
                    if (dt > A.d_min_a_dif) {
                        const float ac = acosf(fminf(fmaxf(dt, -1.0f), 1.0f));
                        const float ov = min_ra[b * 256] - (float)((double)ac * abr) + 1.0f;
                        if (ov > 0.0f) s_add += fminf(max_add, ov) * A.correction_weight;
                    }
                }
            }
            amount += (double)fminf(s_add + ma, 1.0f);
        }
        const double av = rte(16383.0 * (1.0 - amount / (double)rc));
        uint32_t *w = B.gbuf + 5 * (size_t)pix;
        w[4] = (w[4] & ~0xFFFFu) | (uint32_t)fmax(av, 0.0);
        done = true;
    }
    end:
        if (done) have = false;
    }
    flush_counters(cnt, B.ctr);
}

// The DE AO generator's state n draws after `s` (the composition of n
// steps of s -> 0x343FD s + 0x269EC3 mod 2^32, by squaring)
__device__ inline uint32_t lcg_jump(uint32_t s, uint32_t n) {
    uint32_t a = 0x343FDu, c = 0x269EC3u, ma = 1u, mc = 0u;
    while (n) {
        if (n & 1u) {
            ma *= a;
            mc = mc * a + c;
        }
        c *= a + 1u;
        a *= a;
        n >>= 1;
    }
    return ma * s + mc;
}

// The random first steps of the DE AO: each object pixel's generator state,
// the row's seed (DeaoRows::row, src/deao.rs:254-262) advanced by one draw
// per ray of every object pixel left of it, as the row's pass draws them in
// order.  One block per row: the object flags counted by ballot, 256 pixels
// a turn.  `rc_y`: the ray count of the rows by y % side (dithering drops
// the straight ray on shifted rows).
struct AoSeedJob {
    int rc_y[3];
    int side;
};

static __global__ __launch_bounds__(256) void k_ao_seed(DevP P, const uint32_t *gbuf, uint32_t *rnd, AoSeedJob J) {
    __shared__ uint32_t s_cnt[4];
    const int r = blockIdx.x;
    const int y = P.row0 + r;
    const uint32_t seed = (uint32_t)(0x24563487LL + ((long long)y + 1) * 0x324594A1LL);
    const uint32_t rc = (uint32_t)J.rc_y[J.side > 1 ? y % J.side : 0];
    const int lane = (int)(threadIdx.x & 63), wv = (int)(threadIdx.x >> 6);
    const unsigned long long below = lane ? (~0ull >> (64 - lane)) : 0ull;
    uint32_t before = 0;
    for (int x0 = 0; x0 < P.width; x0 += 256) {
        const int x = x0 + (int)threadIdx.x;
        const size_t pix = (size_t)r * P.width + x;
        bool obj = false;
        if (x < P.width) {
            const uint32_t *w = gbuf + 5 * pix;
            obj = (w[2] >> 16) < 32768u && (w[3] & 0xFFFF) < 32768u;
        }
        const unsigned long long m = __ballot(obj);
        if (lane == 0) s_cnt[wv] = (uint32_t)__popcll(m);
        __syncthreads();
        uint32_t off = before + (uint32_t)__popcll(m & below);
        for (int i = 0; i < wv; i++) off += s_cnt[i];
        if (obj) rnd[pix] = lcg_jump(seed, rc * off);
        before += s_cnt[0] + s_cnt[1] + s_cnt[2] + s_cnt[3];
        __syncthreads();
    }
}

// de_at (src/calc.rs:2695-2703) at a list of points: the f64 DE in world
// units and the iteration count
template <int KIND>
__global__ void k_de_at(DevP P, const double *pts, long long n, double *de, int32_t *its) {
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const De64 r = de64<KIND, DEM_NUMERIC>(P, pts[3 * i], pts[3 * i + 1], pts[3 * i + 2], P.max_it, 0, P.de_stop);
        if (de) de[i] = r.de * P.step_width;
        if (its) its[i] = r.it_result;
    }
}

// The general kind's launches (march_gen.hip); `dem` is DEM_NUMERIC or
// DEM_BOX, `f32` is always false for it (no FP32 iterate).
void launch_gen_finish(int grid, hipStream_t s, const DevP &d, const Bufs &b, int all, int dem);
void launch_gen_normals(int grid, hipStream_t s, const DevP &d, const Bufs &b, int dem);
void launch_gen_normals_ext(int grid, hipStream_t s, const DevP &d, const Bufs &b, int dem);
void launch_gen_shadow(int grid, hipStream_t s, const DevP &d, const Bufs &b, const ShadowJob &j, int dem);
void launch_gen_deao(int grid, hipStream_t s, const DevP &d, const Bufs &b, const AoRays &r, int dem);
void launch_gen_de_at(int grid, hipStream_t s, const DevP &d, const double *pts, long long n, double *de,
                      int32_t *its);
void launch_gen_rays(int grid, hipStream_t s, const DevP &d, const mb3d_hip_ray_task *tasks, long long n,
                     int depth_sqrt, mb3d_hip_ray_hit *hits, unsigned long long *ctr, int dem);
void launch_gen_ray_post(int grid, hipStream_t s, const DevP &d, const mb3d_hip_ray_task *tasks, long long n,
                         const RayPost &j, mb3d_hip_ray_hit *hits, unsigned long long *ctr, int dem);

}  // namespace mb3d
