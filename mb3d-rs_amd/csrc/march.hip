// march.hip -- the C ABI of the HIP backend (include/mb3d_hip.h): buffers,
// launches and plans of the DE ray marcher's passes (march_kernels.h), and
// the FP32 kinds' kernel instantiations.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstddef>
#include <cstdlib>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/mb3d_hip.h"
#include "host_common.h"
#include "paint_common.h"
#include "march_kernels.h"
#include "mc_kernel.h"

namespace mb3d {

// clears the counter words (a kernel, not a memset node: the batch plans
// are captured into HIP graphs)
__global__ void k_zero(unsigned long long *p, int n) {
    for (int i = (int)threadIdx.x; i < n; i += (int)blockDim.x) p[i] = 0ull;
}

// the rays of the rectangle's pixels as the network takes them: start
// point and unit direction (f32), and the length of one march unit
__global__ void k_rays(DevP P, float *rays, float *step_len) {
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < P.npix; i += stride) {
        const int y = P.row0 + (int)(i / P.width), x = (int)(i % P.width);
        double c[3], v[3];
        pixel_ray(P, x, y, c, v);
        const double l = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        for (int k = 0; k < 3; k++) {
            rays[6 * i + k] = (float)c[k];
            rays[6 * i + 3 + k] = (float)(v[k] / l);
        }
        step_len[i] = (float)l;
    }
}

// the warm start's proposal buffer behind `base` (8-byte aligned) for
// `items` work items; returns its size in bytes
size_t learn_bufs(void *base, long long items, Bufs *b) {
    if (b) b->prop = (float2 *)base;
    return (size_t)items * sizeof(float2) + 256;
}

// The lists the kernels append to in chunks of 64 (hand-offs, refined
// surface points, hits) hold the items and up to 63 empty slots per wave
// that appended: the capacity for `items` with the largest grid used (the
// march's waves_per_cu, the f64 finish's 32 waves per CU).
size_t list_capacity(long long items, int cus, int wpc) {
    return (size_t)items + 64 * (size_t)cus * (size_t)std::max(wpc, 32);
}

// the surface pixels of a G-buffer marched elsewhere
__global__ void k_collect_hits(DevP P, Bufs B) {
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < P.npix; i += stride) {
        const uint32_t *w = B.gbuf + 5 * (size_t)i;
        const bool surf = (w[2] >> 16) < 32768u && (w[3] & 0xFFFF) < 32768u;
        const long long s = wave_append(surf, &B.ctr[C_HITS]);
        if (surf) B.hits[s] = (int32_t)i;
    }
}

// ---------------------------------------------------------------------------
// host side

thread_local std::string t_error;
thread_local mb3d_hip_counters t_counters;

struct Ctx {
    std::mutex mu;
    int device = -1;
    int cus = 0;
    hipStream_t stream = nullptr;
    // ev[0] start, ev[4] after the ray tables, ev[5] after the cone rays,
    // ev[6] after the network, ev[3] after the FP32 march, ev[7] after the
    // finish, ev[1] after the normals, ev[2] after the post passes
    hipEvent_t ev[8] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    // grow-only device buffers
    void *d_out = nullptr;
    size_t out_bytes = 0;
    void *d_work = nullptr;
    size_t work_bytes = 0;
    unsigned long long *d_ctr = nullptr;
    // the device address of `scratch` (hipGetSymbolAddress at init)
    void *scratch = nullptr;
    // the DE AO's ray sets of the last post job (the source of its copy)
    AoSet ao_sets[9];
};
Ctx g_ctx;

// The no-allocation march's scratch (mb3d_hip_march_noalloc): its hand-off
// list, surface records, hit list, ray tables, cone results and counters,
// carved per call.  Each device that runs it holds the array until the
// process ends; host code uses its device address (Ctx::scratch), never
// this name.  Calls that overlap on one device race on it.
constexpr size_t kScratchBytes = (size_t)256 << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

int set_error(int code, const char *msg) {
    t_error = msg;
    return code;
}

std::mutex &lib_mutex() { return g_ctx.mu; }
hipStream_t lib_stream() { return g_ctx.device >= 0 ? g_ctx.stream : nullptr; }
int lib_cus() { return g_ctx.cus; }
int lib_device() { return g_ctx.device; }
mb3d_hip_counters &thread_counters() { return t_counters; }

#define HIP_TRY(expr)                                                                             \
    do {                                                                                          \
        hipError_t e_ = (expr);                                                                   \
        if (e_ != hipSuccess) {                                                                   \
            char b_[256];                                                                         \
            std::snprintf(b_, sizeof b_, "%s: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, __LINE__); \
            return set_error(MB3D_HIP_ERR_DEVICE, b_);                                            \
        }                                                                                         \
    } while (0)

int kind_of(const mb3d_hip_params &p);
bool ext_modes(const mb3d_hip_params &p);

std::string unsupported_reason(const mb3d_hip_params *p) {
    if (!p) return "no parameters";
    if (p->abi_version != MB3D_HIP_ABI_VERSION) return "parameter ABI version mismatch";
    if (p->n_slots < 1 || p->n_slots > MB3D_HIP_MAX_SLOTS) return "bad hybrid slot count";
    for (int i = 0; i < p->n_slots; i++) {
        if (p->slot_kind[i] < MB3D_HIP_FORMULA_INTPOW || p->slot_kind[i] > MB3D_HIP_FORMULA_AEXIONC)
            return "formula not on this backend";
        const bool second = p->decomb && i > p->end_to;
        if ((p->slot_ade[i] != 0) != ((second ? p->is_custom_de2 : p->is_custom_de) != 0))
            return "a slot's analytic DE flag differs from its part's";
    }
    if (p->decomb) {
        if (p->decomb < 1 || p->decomb > 5) return "this DE combination (not on this backend)";
        if (p->start2 <= p->end_to || p->end2 < p->start2 || p->end2 >= p->n_slots || p->repeat2 < p->start2 ||
            p->repeat2 > p->end2)
            return "the DE combination's second part does not fit the hybrid";
        if (!(p->de_option2 == 0 || p->de_option2 == 4 || p->de_option2 == 11)) return "this DE option of the second part";
        if ((p->four_d2 != 0) != (p->de_option2 == 4)) return "the second part's 4D mode does not fit its DE option";
        if (p->ipol) return "a DE combination of an interpolation hybrid";
        if (p->decomb >= 4 && !(p->decomb_smooth > 0.0)) return "the smooth combination's distance";
    }
    if (!(p->de_option == 0 || p->de_option == 4 || p->de_option == 11 || p->de_option == 32))
        return "this DE option (not on this backend)";
    if ((p->four_d != 0) != (p->de_option >= 4 && p->de_option <= 6)) return "the 4D mode does not fit the DE option";
    if (p->de_option == 32 && kind_of(*p) == K_GEN)
        return "the analytic power-8 DE with another power or in Julia mode (not on this backend)";
    if (p->decomb && (p->color_option != 0 || !p->normals_on_de || p->color_on_it != 0))
        return "a colour option, colour on an iteration or normals on the smooth iteration with a DE combination "
               "(not on this backend)";
    return "";
}

// The secondary rays and the Monte Carlo batches keep the plain normal
// stencil and the orbit-trap colour: scenes with the general surface modes
// are refused there.
std::string rays_unsupported_reason(const mb3d_hip_params *p) {
    std::string why = unsupported_reason(p);
    if (!why.empty()) return why;
    if (ext_modes(*p))
        return "smooth normals, normals on the smooth iteration or colour options in ray batches (not on this "
               "backend)";
    return "";
}

// The normals' stencil in FP32 (the analytic DEs): when its offset (0.15 of
// the DE stop, src/calc.rs:1817-1818) spans at least 128 FP32 ulps of the
// scene's coordinates, so that the rounding of the stencil's points moves a
// normal by well under the 0.5 degree bar; f64 stencils otherwise (closer
// views).
bool fp32_stencil_ok(const mb3d_hip_params &p) {
    const double coord = std::max({std::fabs(p.ystart[0]), std::fabs(p.ystart[1]), std::fabs(p.ystart[2])}) +
                         p.step_width * (p.width + p.height) + 2.0;
    const double noffset = 0.15 * std::fmin((double)p.de_stop, 1.0) * p.step_width;
    return noffset >= 128.0 * coord * 1.1920928955078125e-07;
}

// needs_double (src/gpu.rs:569-576, 1019-1040): the FP32 iterate cannot
// resolve the numeric gradient offset or the pixel step
bool needs_f64(const mb3d_hip_params &p) {
    if (kind_of(p) == K_GEN) return true;  // no FP32 iterate for these formulas
    const double coord = std::max({std::fabs(p.ystart[0]), std::fabs(p.ystart[1]), std::fabs(p.ystart[2])}) +
                         p.step_width * (p.width + p.height) + 2.0;
    const double eps = 1.1920928955078125e-07;
    const double precision = (double)p.de_offset / (coord * eps);
    const double step_precision = p.step_width / (coord * eps);
    const bool numeric = !p.is_custom_de;
    // beyond the reference's rule: a numeric DE's finite differences at
    // de_offset keep too few FP32 bits for the stated tolerance (the FP32
    // march's trajectory drifts; ~0.6 % of a 1080p Mandelbulb's depths
    // leave 1e-4), so a numeric DE marches in FP64 unless FP32 is asked for
    (void)precision;
    return numeric || step_precision < 64.0;
}

// the default of mb3d_hip_options::refill_min: a wave refills once a
// quarter of its lanes are idle (the 1080p bulb's 48-frame plan: 2733
// Mpix/s against 2372 when every turn with an idle lane refills)
constexpr int kRefillMin = 16;

// tile sides: powers of two from 1 to 64, at most 4096 items a tile
bool tile_ok(const mb3d_hip_options &o) {
    auto pow2 = [](int v) { return v == 0 || (v > 0 && v <= 64 && (v & (v - 1)) == 0); };
    return pow2(o.tile_w) && pow2(o.tile_h) && o.tile_w * o.tile_h <= 4096;
}

// the cone gate asks for the network and the cone rays, and a probability
bool gate_ok(const mb3d_hip_options &o) {
    if (o.mlp_cone_gate == 0.0f) return true;
    return o.mlp_cone_gate > 0.0f && o.mlp_cone_gate < 1.0f && o.mlp != nullptr && o.mlp_cone != 0;
}

// entries of a frame's ray tables (k_ray_tab): its columns, then its rows
long long ray_tab_entries(const DevP &d) { return (long long)d.width + d.rows; }

void fill_devp(const mb3d_hip_params &p, const mb3d_hip_options &o, int row0, int rows, DevP &d) {
    std::memset(&d, 0, sizeof d);
    d.width = p.width;
    d.height = p.height;
    d.row0 = row0;
    d.rows = rows;
    d.n_slots = p.n_slots;
    d.end_to = p.end_to;
    d.repeat_from = p.repeat_from;
    d.max_it = p.max_it;
    d.min_it = p.min_it;
    d.de_option = p.de_option;
    d.is_custom_de = p.is_custom_de;
    d.de_add_steps = p.de_add_steps;
    d.dfog_on_it = p.dfog_on_it;
    d.first_step_random = p.first_step_random;
    d.step_cap = p.step_cap;
    d.four_d = p.four_d;
    d.do_julia = p.do_julia;
    d.ipol = p.ipol;
    d.ipol_w[0] = p.ipol_w[0];
    d.ipol_w[1] = p.ipol_w[1];
    d.decomb = p.decomb;
    d.start2 = p.start2;
    d.end2 = p.end2;
    d.repeat2 = p.repeat2;
    d.max_it2 = p.max_it2;
    d.de_option2 = p.de_option2;
    d.custom2 = p.is_custom_de2;
    d.four_d2 = p.four_d2;
    d.d_de_scale2 = p.d_de_scale2;
    d.smooth = p.decomb_smooth;
    d.placed = p.placed;
    for (int k = 0; k < 3; k++) {
        d.place_off[k] = p.place_offset[k];
        for (int q = 0; q < 3; q++) d.place_m[k][q] = p.place_m[k][q];
    }
    d.place_scale = p.place_scale;
    for (int k = 0; k < 4; k++) d.ju[k] = p.julia_c[k];
    for (int i = 0; i < 4; i++)
        for (int k = 0; k < 3; k++) d.sm4[3 * i + k] = (float)p.smatrix4[i][k];
    for (int i = 0; i < MB3D_HIP_MAX_SLOTS; i++) {
        d.slot_kind[i] = p.slot_kind[i];
        d.slot_iters[i] = std::max(p.slot_iters[i], 0);
        // the f64 options (DevP::opt, then opt2) and the integer argument
        double o[6] = {0, 0, 0, 0, 0, 0};
        const double *so = p.slot_opt[i];
        switch (p.slot_kind[i]) {
            case MB3D_HIP_FORMULA_INTPOW:
                o[0] = p.slot_z_mul[i];
                d.slot_iarg[i] = p.slot_power[i];
                break;
            case MB3D_HIP_FORMULA_AMAZINGBOX:
                o[0] = p.slot_scale[i];
                o[1] = p.slot_min_r[i];
                o[2] = p.slot_fold[i];
                break;
            case MB3D_HIP_FORMULA_FOLDINGINTPOW:
                o[0] = so[1];
                o[1] = so[2];
                d.slot_iarg[i] = p.slot_power[i];
                break;
            case MB3D_HIP_FORMULA_AEXIONC:
                o[0] = so[0];
                o[1] = so[1];
                o[2] = so[4];
                o[3] = so[5];
                d.slot_iarg[i] = (so[2] != 0.0 ? 1 : 0) | (so[3] != 0.0 ? 2 : 0) | (so[6] != 0.0 ? 4 : 0) |
                                 ((int)so[7] << 8);
                break;
            default:  // Real Power, Quaternion, Tricorn, Bulbox: options() in order
                for (int k = 0; k < 6; k++) o[k] = so[k];
                break;
        }
        for (int k = 0; k < 3; k++) {
            d.opt[i][k] = o[k];
            d.opt2[i][k] = o[3 + k];
        }
        d.slot_ade[i] = p.slot_ade[i];
        d.slot_unc[i] = p.slot_uncounted[i] != 0;
        d.fz_mul[i] = (float)p.slot_z_mul[i];
        d.fscale[i] = (float)p.slot_scale[i];
        const double m = std::fmax(p.slot_min_r[i], 1e-40);
        d.fmin_r2[i] = (float)(m * m);
        d.fscale_min_r2[i] = (float)(p.slot_scale[i] / (m * m));
        d.ffold[i] = (float)p.slot_fold[i];
        d.fhln[i] = p.fhln[i];
    }
    d.d_de_scale = p.d_de_scale;
    d.de_stop = p.de_stop;
    d.z_step_div = p.z_step_div;
    d.mh04zsd = p.mh04zsd;
    d.de_stop_factor = p.de_stop_factor;
    d.ln_rstop = p.ln_rstop;
    d.de_offset = p.de_offset;
    d.de_offset006 = p.de_offset006;
    d.mcts_m = p.mcts_m;
    d.ms_de_sub = p.ms_de_sub;
    d.d_col_plus = p.d_col_plus;
    d.col_var = p.col_var_de_stop_mul;
    d.fovx_off = p.fovx_off;
    d.fovx_mul = p.fovx_mul;
    // (0.05: the FP32 DE's error near the surface is ~1e-5 of it at the
    // zooms that take the FP32 march; 0.25 and 0.02 gave the same accuracy
    // against the oracle on the 1080p bulb, and 0.05 takes 6 % less time)
    d.guard = 1.0f + (o.guard_band > 0.0f ? o.guard_band : 0.02f);
    d.f_d_rstop = (float)p.d_rstop;
    d.f_rstop3d = (float)p.rstop3d;
    d.d_rstop = p.d_rstop;
    d.rstop3d = p.rstop3d;
    d.fov_y = p.fov_y;
    d.zcorr = p.zcorr;
    d.zc_mul = p.zc_mul;
    d.zend = p.zend;
    d.step_width = p.step_width;
    std::memcpy(d.vg, p.vgrads, sizeof d.vg);
    std::memcpy(d.ystart, p.ystart, sizeof d.ystart);
    d.it_split = o.it_split > 0 ? o.it_split : 0;
    d.tile_w = o.tile_w > 0 ? o.tile_w : 8;
    d.tile_h = o.tile_h > 0 ? o.tile_h : 8;
    d.tw_shift = __builtin_ctz((unsigned)d.tile_w);
    d.tn_shift = d.tw_shift + __builtin_ctz((unsigned)d.tile_h);
    d.refill_min = o.refill_min > 0 ? std::min(o.refill_min, 64) : kRefillMin;
    d.rt_base = 0;
    d.cone_mode = std::min(std::max(o.mlp_cone, 0), 2);
    d.sm_normals = std::min(std::max(p.sm_normals, 0), 8);
    d.normals_on_de = p.normals_on_de != 0;
    d.color_option = std::min(std::max(p.color_option, 0), 5);
    d.color_on_it = std::min(std::max(p.color_on_it, 0), 255);
    d.mct_color_mul = p.mct_color_mul;
    d.tiles_x = (p.width + d.tile_w - 1) / d.tile_w;
    const long long tiles_y = (rows + d.tile_h - 1) / d.tile_h;
    d.npix = (long long)p.width * rows;
    d.nslots_q = (long long)d.tiles_x * tiles_y * d.tile_w * d.tile_h;
}

// build_rays (src/deao.rs:75-105); quality 4 is the 8-ray set (extension):
// the quality-1 geometry with 7 rays on its ring
void build_rays(int quality, float max_len, AoRays &r, int row_count[4] = nullptr) {
    std::memset(&r, 0, sizeof r);
    if (row_count) std::fill(row_count, row_count + 4, 0);
    const double TAU = 6.283185307179586;
    auto push = [&](double ya, double za) {
        const double sy = std::sin(ya), cy = std::cos(ya), sz = std::sin(za), cz = std::cos(za);
        r.dir[r.count][0] = (float)(sy * cz);
        r.dir[r.count][1] = (float)(-sy * sz);
        r.dir[r.count][2] = (float)(-cy);
        r.count++;
    };
    const int q = std::min(std::max(quality, 0), 4);
    r.max_len = max_len;
    if (q == 0) {
        const double abr = 60.0 * (3.141592653589793 / 180.0);
        for (int x = 0; x < 3; x++) push(0.5 * abr, x * TAU / 3.0);
        r.d_step_mul = 1.8f;
        r.d_min_a_dif = -1.0f;
        r.correction_weight = 0.3f;
        return;
    }
    const int qg = q == 4 ? 1 : q;
    push(0.0, 0.0);
    const double abr = 1.5707963267948966 / ((double)qg + 0.9);
    for (int y = 1; y <= qg; y++) {
        const double dt1 = y * abr;
        int n = (int)std::nearbyint(std::sin(dt1) * TAU / abr);
        if (q == 4) n = 7;
        if (row_count) row_count[y] = n;
        for (int x = 0; x < n; x++) push(dt1, x * TAU / n);
    }
    r.d_step_mul = (float)(1.0 + std::sin(abr));
    r.d_min_a_dif = (float)std::cos(abr * 1.2);
    r.correction_weight = qg == 1 ? 0.2f : 0.1666f;
}

// The DE AO's ray sets per dither cell (src/deao.rs:318-342): for the cell
// (x % side, y % side) the rings shifted by half a ring and half a gap per
// side step, the straight ray dropped on the shifted rows; side = dither +
// 1.  Without dithering the one set of build_rays.  Quality 0..3.
void build_ao_sets(int quality, int dither, AoSet *sets) {
    AoRays base;
    int row_count[4];
    build_rays(quality, 1.0f, base, row_count);
    std::memset(sets, 0, 9 * sizeof(AoSet));
    if (dither <= 0) {
        std::memcpy(sets[0].dir, base.dir, sizeof base.dir);
        sets[0].count = base.count;
        return;
    }
    const double TAU = 6.283185307179586;
    const int side = dither + 1;
    quality = std::min(std::max(quality, 0), 3);
    for (int yc = 0; yc < side; yc++)
        for (int xc = 0; xc < side; xc++) {
            AoSet &o = sets[yc * side + xc];
            const double dt1 = (double)yc * 0.5 / (double)dither;
            const double dt2 = (double)xc * 0.5 / (double)dither;
            auto push = [&](double ya, double za) {
                const double sy = std::sin(ya), cy = std::cos(ya), sz = std::sin(za), cz = std::cos(za);
                o.dir[o.count][0] = (float)(sy * cz);
                o.dir[o.count][1] = (float)(-sy * sz);
                o.dir[o.count][2] = (float)(-cy);
                o.count++;
            };
            if (quality == 0) {
                for (int i = 0; i < 3; i++)
                    push((dt1 + 0.5) * (50.0 * (3.141592653589793 / 180.0)), ((double)i + dt2) * TAU / 3.0);
                continue;
            }
            const double abr = 1.5707963267948966 / ((double)quality + 0.9);
            if (!(dt1 > 0.1)) push(0.0, 0.0);
            for (int rr = 1; rr <= quality; rr++) {
                const int n = row_count[rr];
                for (int i = 0; i < n; i++) push(abr * ((double)rr + dt1 - 0.25), ((double)i + dt2) * TAU / (double)n);
            }
        }
}

int ensure(void **buf, size_t *have, size_t need) {
    if (*have >= need) return 0;
    if (*buf) (void)hipFree(*buf);
    *buf = nullptr;
    *have = 0;
    size_t n = std::max(need, (size_t)1 << 20);
    HIP_TRY(hipMalloc(buf, n));
    *have = n;
    return 0;
}

void *lib_work(size_t bytes, int *status) {
    *status = ensure(&g_ctx.d_work, &g_ctx.work_bytes, bytes);
    return *status ? nullptr : g_ctx.d_work;
}

int kind_of(const mb3d_hip_params &p) {
    bool p8 = false, box = false;
    if (p.four_d || p.do_julia || p.ipol || p.decomb) return K_GEN;
    for (int i = 0; i < p.n_slots; i++) {
        if (p.slot_iters[i] == 0) continue;
        if (p.slot_uncounted[i]) return K_GEN;
        const int k = p.slot_kind[i];
        if (k == MB3D_HIP_FORMULA_INTPOW && p.slot_power[i] == 8) p8 = true;
        else if (k == MB3D_HIP_FORMULA_AMAZINGBOX) box = true;
        else return K_GEN;
    }
    if (p8 && box) return K_HYB;
    return box ? K_BOX : K_P8;
}

int dem_of(const mb3d_hip_params &p) {
    if (p.de_option == 11) return DEM_BOX;
    if (p.de_option == 32) return DEM_P8;
    return DEM_NUMERIC;
}

// the warm start's jumps are on (hit or miss jumps: the per-item proposals
// and the learned march); the network may be on for the cone's gate alone
bool jumps_on(const MlpDev &M) { return M.enabled && (M.conf_hit <= 1.0f || M.conf_miss >= 0.0f); }

// the network's pass: per work item (the warm start's proposals), or per
// tile (`tiles`: the cone pass's gate)
template <bool BATCH>
void launch_propose(hipStream_t s, const DevP &d, const DevP *tab, int n, const MlpDev &M, const double2 *rtab,
                    const ConeRes *cone, float2 *prop, unsigned long long *ctr, bool tiles = false) {
    const long long groups = ((d.nslots_q >> (tiles ? d.tn_shift : 0)) + 63) / 64;
    const int grid = (int)std::max<long long>(1, std::min<long long>((groups + 3) / 4, (long long)g_ctx.cus * 16));
    hipLaunchKernelGGL((k_mlp_propose<BATCH>), dim3(grid), dim3(256), 0, s, d, tab, n, M, rtab, cone, prop, ctr,
                       tiles ? 1 : 0);
}

// the cone rays; with the network's gate (M.gate) its per-tile pass first,
// into `prop`
template <int KIND, int DEM, bool BATCH>
void launch_cone(hipStream_t s, const DevP &d, const DevP *tab, int n, const double2 *rtab, ConeRes *out,
                 unsigned long long *ctr, const MlpDev &M, float2 *prop) {
    if constexpr (KIND == K_GEN) return;  // the FP32 march's kinds only
    else {
        const bool gate = M.enabled && M.gate && prop;
        if (gate) launch_propose<BATCH>(s, d, tab, n, M, rtab, nullptr, prop, ctr, true);
        const long long tiles = d.nslots_q >> d.tn_shift;
        const int grid = (int)std::max<long long>(1, std::min<long long>((tiles + 255) / 256, (long long)g_ctx.cus * 8));
        hipLaunchKernelGGL((k_cone<KIND, DEM, BATCH>), dim3(grid), dim3(256), 0, s, d, tab, n, rtab, out, ctr,
                           gate ? (const float2 *)prop : nullptr, M.gate_logit);
    }
}

// the ray tables of one frame (tab NULL: `d`) or of a batch's n frames
void launch_ray_tab(hipStream_t s, const DevP &d, const DevP *tab, int n, double2 *rtab) {
    hipLaunchKernelGGL(k_ray_tab, dim3(8, tab ? n : 1), dim3(256), 0, s, tab, d, rtab);
}

// the FP32 march; with the cone rays k_cone_bg first: the background tiles'
// pixels, and the list of tiles the march takes
template <int KIND, int DEM, bool BATCH = false>
void launch_march(int grid, hipStream_t s, const DevP &d, const Bufs &b, const MlpDev &M, const DevP *tab = nullptr,
                  int n = 1) {
    const size_t lds = BATCH ? (size_t)n * DEVP_CORE : 0;
    if constexpr (KIND == K_GEN) {
        return;  // FP64 only: the hosts never ask for it
    } else {
        if (d.cone_mode) {
            const long long groups = ((d.nslots_q >> d.tn_shift) + 63) / 64;  // 64 tiles a wave
            const int gbg =
                (int)std::max<long long>(1, std::min<long long>((groups + 3) / 4, (long long)g_ctx.cus * 8));
            hipLaunchKernelGGL((k_cone_bg<KIND, DEM, BATCH>), dim3(gbg), dim3(256), lds, s, d, b, tab, n);
        }
        if (jumps_on(M))
            hipLaunchKernelGGL((k_march32<KIND, DEM, BATCH, true>), dim3(grid), dim3(256), lds, s, d, b, tab, n, M);
        else
            hipLaunchKernelGGL((k_march32<KIND, DEM, BATCH, false>), dim3(grid), dim3(256), lds, s, d, b, tab, n, M);
    }
}
// (the general kind's kernels are instantiated in march_gen.hip)
template <int KIND, int DEM, bool BATCH = false>
void launch_finish(int grid, hipStream_t s, const DevP &d, const Bufs &b, int all, const DevP *tab = nullptr,
                   int n = 1) {
    const size_t lds = BATCH ? (size_t)n * DEVP_CORE : 0;
    if constexpr (KIND == K_GEN) {
        if (!BATCH) launch_gen_finish(grid, s, d, b, all, DEM);  // FP64 only, never batched
    } else if (BATCH || !all) {
        // the FP32 march's finish: the fast f64 forms
        hipLaunchKernelGGL((k_finish64<KIND, DEM, BATCH, true>), dim3(grid), dim3(256), lds, s, d, b, all, tab, n);
    } else {
        // the FP64 mode: the reference's operation order
        hipLaunchKernelGGL((k_finish64<KIND, DEM, false, false>), dim3(grid), dim3(256), lds, s, d, b, all, tab, n);
    }
}
// the surface modes beyond the plain stencil (k_normals_ext): smooth
// normals, normals on the smooth iteration, colour options, colour on an
// iteration
bool ext_modes(const mb3d_hip_params &p) {
    return p.sm_normals != 0 || !p.normals_on_de || p.color_option != 0 || p.color_on_it != 0;
}

// the normals pass after the finish (N32: FP32 offset DEs, analytic DEs
// only; ext: k_normals_ext for the general surface modes)
template <int KIND, int DEM, bool BATCH = false>
void launch_normals(int grid, hipStream_t s, const DevP &d, const Bufs &b, bool f32, const DevP *tab = nullptr,
                    int n = 1, bool ext = false) {
    const size_t lds = BATCH ? (size_t)n * DEVP_CORE : 0;
    if constexpr (KIND == K_GEN) {
        if (!BATCH) {
            if (ext)
                launch_gen_normals_ext(grid, s, d, b, DEM);
            else
                launch_gen_normals(grid, s, d, b, DEM);
        }
    } else if (ext) {
        hipLaunchKernelGGL((k_normals_ext<KIND, DEM, BATCH>), dim3(grid), dim3(256), lds, s, d, b, tab, n);
    } else if constexpr (DEM != DEM_NUMERIC) {
        if (f32)
            hipLaunchKernelGGL((k_normals<KIND, DEM, true, BATCH>), dim3(grid), dim3(256), lds, s, d, b, tab, n);
        else
            hipLaunchKernelGGL((k_normals<KIND, DEM, false, BATCH>), dim3(grid), dim3(256), lds, s, d, b, tab, n);
    } else {
        hipLaunchKernelGGL((k_normals<KIND, DEM, false, BATCH>), dim3(grid), dim3(256), lds, s, d, b, tab, n);
    }
}
template <int KIND, int DEM, bool F32>
void launch_shadow(int grid, hipStream_t s, const DevP &d, const Bufs &b, const ShadowJob &j) {
    if constexpr (KIND == K_GEN) {
        if (!F32) launch_gen_shadow(grid, s, d, b, j, DEM);
    } else {
        hipLaunchKernelGGL((k_shadow<KIND, DEM, F32>), dim3(grid), dim3(256), 0, s, d, b, j);
    }
}
template <int KIND, int DEM, bool F32>
void launch_deao(int grid, hipStream_t s, const DevP &d, const Bufs &b, const AoRays &r) {
    if constexpr (KIND == K_GEN) {
        if (!F32) launch_gen_deao(grid, s, d, b, r, DEM);
    } else {
        hipLaunchKernelGGL((k_deao<KIND, DEM, F32>), dim3(grid), dim3(256), 0, s, d, b, r);
    }
}

template <int KIND>
void launch_de_at(int grid, hipStream_t s, const DevP &d, const double *pts, long long n, double *de, int32_t *its) {
    if constexpr (KIND == K_GEN)
        launch_gen_de_at(grid, s, d, pts, n, de, its);
    else
        hipLaunchKernelGGL((k_de_at<KIND>), dim3(grid), dim3(256), 0, s, d, pts, n, de, its);
}

template <int KIND, int DEM>
void launch_rays(int grid, hipStream_t s, const DevP &d, const mb3d_hip_ray_task *tasks, long long n, int depth_sqrt,
                 mb3d_hip_ray_hit *hits, unsigned long long *ctr) {
    if constexpr (KIND == K_GEN)
        launch_gen_rays(grid, s, d, tasks, n, depth_sqrt, hits, ctr, DEM);
    else
        hipLaunchKernelGGL((k_rays<KIND, DEM>), dim3(grid), dim3(256), 0, s, d, tasks, n, depth_sqrt, hits, ctr);
}

template <int KIND, int DEM>
void launch_ray_post(int grid, hipStream_t s, const DevP &d, const mb3d_hip_ray_task *tasks, long long n,
                     const RayPost &j, mb3d_hip_ray_hit *hits, unsigned long long *ctr) {
    if constexpr (KIND == K_GEN)
        launch_gen_ray_post(grid, s, d, tasks, n, j, hits, ctr, DEM);
    else
        hipLaunchKernelGGL((k_ray_post<KIND, DEM>), dim3(grid), dim3(256), 0, s, d, tasks, n, j, hits, ctr);
}

template <int KIND, int DEM>
void launch_mc(int grid, hipStream_t s, const DevP &d, const mb3d_hip_mc_task *tasks, long long n, int secant,
               mb3d_hip_mc_res *res, unsigned long long *ctr) {
    if constexpr (KIND == K_GEN)
        launch_gen_mc(grid, s, d, tasks, n, secant, res, ctr, DEM);
    else
        hipLaunchKernelGGL((k_mc<KIND, DEM>), dim3(grid), dim3(256), 0, s, d, tasks, n, secant, res, ctr);
}

// The hit pass's tables from the options: the lights, and per AO quality
// 0..3 deao_1pos's ray set, step factor, overlap weight and the overlap term
// acos(d_k . d_k2) abr of each pair near enough (src/reflect.rs:954-1090).
int ray_post_of(const mb3d_hip_ray_opts &o, RayPost &j) {
    std::memset(&j, 0, sizeof j);
    if (o.n_lights < 0 || o.n_lights > MB3D_HIP_MAX_LIGHTS) return set_error(MB3D_HIP_ERR_INVALID, "bad light count");
    if (o.soft_radius > 0.0f && o.n_lights < 1) return set_error(MB3D_HIP_ERR_INVALID, "a soft shadow needs a light");
    if (o.ao && o.ao_first_random)
        return set_error(MB3D_HIP_UNSUPPORTED, "a random first AO step at ray hits (not on this backend)");
    j.n_lights = o.n_lights;
    j.soft = o.soft_radius > 0.0f;
    j.ao = o.ao != 0;
    j.soft_radius = o.soft_radius;
    j.hs_mul = o.hs_mul;
    j.ao_max_len = o.ao_max_len;
    for (int i = 0; i < o.n_lights; i++) {
        if (o.light_idx[i] < 0 || o.light_idx[i] >= MB3D_HIP_MAX_LIGHTS)
            return set_error(MB3D_HIP_ERR_INVALID, "bad light index");
        j.bit[i] = o.light_idx[i];
        j.positional[i] = o.light_positional[i] != 0;
        j.visible[i] = o.light_visible[i];
        j.lmax[i] = o.light_lmax[i];
        for (int k = 0; k < 3; k++) j.vec[i][k] = o.light_vec[i][k];
    }
    const double tau = 6.283185307179586;
    auto ray_dir = [](double ya, double za, float out[3]) {
        out[0] = (float)(std::sin(ya) * std::cos(za));
        out[1] = (float)(-std::sin(ya) * std::sin(za));
        out[2] = (float)(-std::cos(ya));
    };
    int nd = 0, pen = 0;
    for (int q = 0; q < 4; q++) {
        j.first[q] = nd;
        float dmin;
        if (q == 0) {
            const double abr = 60.0 * (3.141592653589793 / 180.0);
            for (int i = 0; i < 3; i++) ray_dir(0.5 * abr, (double)i * tau / 3.0, j.dirs[nd++]);
            j.d_step_mul[q] = 1.8;
            dmin = -1.0f;
            j.corr_w[q] = 0.3f;
        } else {
            const double abr = 1.5707963267948966 / ((double)q + 0.9);
            ray_dir(0.0, 0.0, j.dirs[nd++]);
            for (int iy = 1; iy <= q; iy++) {
                const int n = (int)std::nearbyint(std::sin(0.25) * tau / abr);
                for (int ix = 0; ix < n; ix++) ray_dir(abr * (double)iy, (double)ix * tau / (double)n, j.dirs[nd++]);
            }
            j.d_step_mul[q] = 1.0 + std::sin(abr);
            dmin = (float)std::cos(abr * 1.2);
            j.corr_w[q] = q == 1 ? 0.2f : 0.1666f;
        }
        const int cnt = nd - j.first[q];
        j.count[q] = cnt;
        j.de_mul[q] = std::sqrt((double)cnt * 0.5);
        const double abr2 = 1.2 / std::asin(std::fmin(std::fmax(1.0 / j.de_mul[q], -1.0), 1.0));
        j.pen_off[q] = pen;
        for (int a = 0; a < cnt; a++)
            for (int b = 0; b < cnt; b++) {
                const float *da = j.dirs[j.first[q] + a], *db = j.dirs[j.first[q] + b];
                const float dt = da[0] * db[0] + da[1] * db[1] + da[2] * db[2];
                const float ac = std::acos(std::fmin(std::fmax(dt, -1.0f), 1.0f));  // f32 acos, as the reference
                j.pen[pen++] = dt > dmin ? (float)((double)ac * abr2) : INFINITY;
            }
    }
    return 0;
}

// dispatch over the (formula kind, DE mode) pairs that exist
template <class F>
void dispatch(int kind, int dem, F &&f) {
    if (kind == K_P8 && dem == DEM_NUMERIC) f(std::integral_constant<int, K_P8>(), std::integral_constant<int, DEM_NUMERIC>());
    else if (kind == K_P8 && dem == DEM_P8) f(std::integral_constant<int, K_P8>(), std::integral_constant<int, DEM_P8>());
    else if (kind == K_BOX && dem == DEM_BOX) f(std::integral_constant<int, K_BOX>(), std::integral_constant<int, DEM_BOX>());
    else if (kind == K_BOX && dem == DEM_NUMERIC) f(std::integral_constant<int, K_BOX>(), std::integral_constant<int, DEM_NUMERIC>());
    else if (kind == K_GEN && dem == DEM_BOX) f(std::integral_constant<int, K_GEN>(), std::integral_constant<int, DEM_BOX>());
    else if (kind == K_GEN) f(std::integral_constant<int, K_GEN>(), std::integral_constant<int, DEM_NUMERIC>());
    else f(std::integral_constant<int, K_HYB>(), std::integral_constant<int, DEM_NUMERIC>());
}

int init_locked(int device) {
    Ctx &c = g_ctx;
    if (c.device == device) return 0;
    if (c.device >= 0) return set_error(MB3D_HIP_ERR_INVALID, "the library is bound to another device");
    int n = 0;
    HIP_TRY(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) return set_error(MB3D_HIP_ERR_DEVICE, "no such HIP device");
    HIP_TRY(hipSetDevice(device));
    HIP_TRY(hipDeviceGetAttribute(&c.cus, hipDeviceAttributeMultiprocessorCount, device));
    HIP_TRY(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
    for (auto &e : c.ev) HIP_TRY(hipEventCreate(&e));
    HIP_TRY(hipMalloc((void **)&c.d_ctr, C_WORDS * sizeof(unsigned long long)));
    HIP_TRY(hipGetSymbolAddress(&c.scratch, HIP_SYMBOL(scratch)));
    c.device = device;
    return 0;
}

struct OutLayout {
    size_t depth, depth64, iters, normals, gbuf, total;
};

OutLayout layout_for(size_t npix) {
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    OutLayout l;
    l.depth = 0;
    l.depth64 = al(l.depth + npix * 4);
    l.iters = al(l.depth64 + npix * 8);
    l.normals = al(l.iters + npix * 2);
    l.gbuf = al(l.normals + npix * 12);
    l.total = al(l.gbuf + npix * 20);
    return l;
}

// `noalloc`: the march takes its buffers from the scratch array (no device
// allocation; device outputs, no post job, no network)
int march_locked(const mb3d_hip_params *params, const mb3d_hip_post_job *post, int row0, int rows,
                 const mb3d_hip_out *out, const mb3d_hip_options *opt_in, mb3d_hip_progress_fn progress,
                 mb3d_hip_cancel_fn cancel, void *user, uint32_t *post_gbuf, int post_on_device,
                 bool noalloc = false) {
    Ctx &c = g_ctx;
    if (c.device < 0) return set_error(MB3D_HIP_ERR_NOT_INIT, "mb3d_hip_init was not called");
    if (!params) return set_error(MB3D_HIP_ERR_INVALID, "params is NULL");
    const mb3d_hip_params &p = *params;
    const std::string why = unsupported_reason(params);
    if (!why.empty()) return set_error(MB3D_HIP_UNSUPPORTED, why.c_str());
    if (p.width <= 0 || p.height <= 0) return set_error(MB3D_HIP_ERR_INVALID, "bad picture size");
    if (row0 < 0 || rows <= 0 || row0 + rows > p.height) return set_error(MB3D_HIP_ERR_INVALID, "rows out of the picture");
    if (post && (post->n_lights < 0 || post->n_lights > MB3D_HIP_MAX_LIGHTS))
        return set_error(MB3D_HIP_ERR_INVALID, "bad light count");
    if (post && post->deao && (post->deao_dither < 0 || post->deao_dither > 2))
        return set_error(MB3D_HIP_ERR_INVALID, "bad DE AO dithering (0, 1 or 2)");
    if (post && post->deao && post->deao_dither && post->deao_quality >= 4)
        return set_error(MB3D_HIP_UNSUPPORTED, "dithering of the 8-ray AO set (quality 4 has no dithered form)");
    mb3d_hip_options o;
    std::memset(&o, 0, sizeof o);
    if (opt_in) o = *opt_in;
    if (!tile_ok(o)) return set_error(MB3D_HIP_ERR_INVALID, "bad tile shape (sides are powers of two up to 64)");
    if (!gate_ok(o))
        return set_error(MB3D_HIP_ERR_INVALID, "mlp_cone_gate is a probability in (0, 1) and needs mlp and mlp_cone");
    if (cancel && cancel(user)) return set_error(MB3D_HIP_CANCELLED, "cancelled");
    // (the K_GEN formulas have no FP32 iterate: FP64 whatever was asked)
    const bool f64 = o.precision == MB3D_HIP_PREC_F64 || (o.precision == MB3D_HIP_PREC_AUTO && needs_f64(p)) ||
                     kind_of(p) == K_GEN;
    DevP d;
    fill_devp(p, o, row0, rows, d);
    const size_t npix = (size_t)d.npix;
    const bool post_only = post_gbuf != nullptr;
    const bool want_shadow = post && post->n_lights > 0;
    const bool want_ao = post && post->deao;

    // buffers: outputs on the device (the caller's or ours), hand-off list, hit list
    const OutLayout L = layout_for(npix);
    const bool dev_out = post_only ? post_on_device != 0 : (out && out->on_device);
    if (noalloc && (!dev_out || post || post_only))
        return set_error(MB3D_HIP_ERR_INVALID, "the no-allocation march writes device outputs (out->on_device = 1), "
                                               "without a post job");
    if (noalloc && o.mlp) return set_error(MB3D_HIP_ERR_INVALID, "the no-allocation march takes no network");
    if (!dev_out) {
        if (int e = ensure(&c.d_out, &c.out_bytes, L.total)) return e;
    }
    MlpDev M;
    mlp_dev(o.mlp, o, M);
    if (f64 || post_only) M.enabled = 0;
    const int wpc = o.waves_per_cu > 0 ? o.waves_per_cu : 16;
    const size_t cap = list_capacity((long long)npix, c.cus, wpc);
    const size_t base_need = (cap * (sizeof(Cand) + sizeof(SurfRec) + sizeof(int32_t)) + 511) & ~(size_t)255;
    const size_t learn_need = M.enabled ? learn_bufs(nullptr, d.nslots_q, nullptr) : 0;
    const size_t rtab_bytes = ((size_t)ray_tab_entries(d) * sizeof(double2) + 255) & ~(size_t)255;
    const bool cone = d.cone_mode != 0 && !f64 && !post_only;
    // (the march's tile list packs a tile index in LIVE_TILE_BITS bits)
    if (cone && (d.nslots_q >> d.tn_shift) >= (1ll << LIVE_TILE_BITS))
        return set_error(MB3D_HIP_ERR_INVALID, "with the cone rays a picture holds at most 2^26 tiles");
    const size_t cone_bytes = cone ? (size_t)(d.nslots_q >> d.tn_shift) * sizeof(ConeRes) : 0;
    const bool ao_rnd = want_ao && post->deao_first_step_random;
    const size_t ao_sets_bytes = want_ao ? (9 * sizeof(AoSet) + 255) & ~(size_t)255 : 0;
    const size_t ao_rnd_bytes = ao_rnd ? (npix * sizeof(uint32_t) + 255) & ~(size_t)255 : 0;
    const size_t ao_off = base_need + learn_need + rtab_bytes + ((cone_bytes + 255) & ~(size_t)255);
    const size_t work_need = ao_off + ao_sets_bytes + ao_rnd_bytes;
    void *work = nullptr;
    unsigned long long *ctr_dev = c.d_ctr;
    if (noalloc) {
        const size_t ctr_off = (work_need + 255) & ~(size_t)255;
        if (ctr_off + C_WORDS * sizeof(unsigned long long) > kScratchBytes) {
            char msg[200];
            std::snprintf(msg, sizeof msg,
                          "the no-allocation march's scratch (%zu MiB) holds about %lld pixels a call: march the "
                          "picture in bands of rows",
                          kScratchBytes >> 20,
                          (long long)((kScratchBytes - 64 * (size_t)c.cus * 32 * (sizeof(Cand) + sizeof(SurfRec) + 4)) /
                                      (sizeof(Cand) + sizeof(SurfRec) + 4 + 32)));
            return set_error(MB3D_HIP_ERR_INVALID, msg);
        }
        work = c.scratch;
        ctr_dev = (unsigned long long *)((char *)c.scratch + ctr_off);
    } else {
        if (int e = ensure(&c.d_work, &c.work_bytes, work_need)) return e;
        work = c.d_work;
    }
    Bufs b;
    std::memset(&b, 0, sizeof b);
    char *ob = (char *)c.d_out;
    if (post_only) {
        b.gbuf = dev_out ? post_gbuf : (uint32_t *)(ob + L.gbuf);
    } else if (dev_out) {
        b.depth = out->depth;
        b.depth64 = out->depth64;
        b.iters = out->iters;
        b.normals = out->normals;
        b.gbuf = out->gbuf;
        if (!b.gbuf && (want_shadow || want_ao)) {
            if (int e = ensure(&c.d_out, &c.out_bytes, L.total)) return e;
            b.gbuf = (uint32_t *)((char *)c.d_out + L.gbuf);
        }
    } else {
        b.depth = out && out->depth ? (float *)(ob + L.depth) : nullptr;
        b.depth64 = out && out->depth64 ? (double *)(ob + L.depth64) : nullptr;
        b.iters = out && out->iters ? (uint16_t *)(ob + L.iters) : nullptr;
        b.normals = out && out->normals ? (float *)(ob + L.normals) : nullptr;
        b.gbuf = (out && out->gbuf) || want_shadow || want_ao ? (uint32_t *)(ob + L.gbuf) : nullptr;
    }
    b.cand = (Cand *)work;
    b.surf = (SurfRec *)((char *)work + cap * sizeof(Cand));
    b.hits = (int32_t *)(b.surf + cap);
    b.ctr = ctr_dev;
    if (M.enabled) learn_bufs((char *)work + base_need, d.nslots_q, &b);
    double2 *rtab = (double2 *)((char *)work + base_need + learn_need);
    b.rtab = rtab;
    b.cone = cone ? (ConeRes *)((char *)rtab + rtab_bytes) : nullptr;
    if (!cone) d.cone_mode = 0;

    if (noalloc)
        hipLaunchKernelGGL(k_zero, dim3(1), dim3(64), 0, c.stream, ctr_dev, (int)C_WORDS);  // (no memset node)
    else
        HIP_TRY(hipMemsetAsync(ctr_dev, 0, C_WORDS * sizeof(unsigned long long), c.stream));
    if (post_only && !dev_out)
        HIP_TRY(hipMemcpyAsync(b.gbuf, post_gbuf, npix * 20, hipMemcpyHostToDevice, c.stream));
    const int kind = kind_of(p), dem = dem_of(p);
    hipStream_t stream = c.stream;
    const int grid_march = std::max(1, c.cus * wpc / 4);
    int grid_flat = std::max(1, c.cus * 4);
    HIP_TRY(hipEventRecord(c.ev[0], c.stream));
    if (!post_only) {
        if (f64) {
            dispatch(kind, dem, [&](auto K, auto D) {
                launch_finish<decltype(K)::value, decltype(D)::value>(grid_flat * 2, stream, d, b, 1);
                launch_normals<decltype(K)::value, decltype(D)::value>(grid_flat, stream, d, b, false, nullptr, 1,
                                                                       ext_modes(p));
            });
        } else {
            launch_ray_tab(stream, d, nullptr, 1, rtab);
            HIP_TRY(hipEventRecord(c.ev[4], c.stream));
            if (cone) {
                dispatch(kind, dem, [&](auto K, auto D) {
                    launch_cone<decltype(K)::value, decltype(D)::value, false>(stream, d, nullptr, 1, rtab, b.cone,
                                                                              ctr_dev, M, b.prop);
                });
            }
            HIP_TRY(hipEventRecord(c.ev[5], c.stream));
            if (jumps_on(M)) launch_propose<false>(stream, d, nullptr, 1, M, rtab, b.cone, b.prop, ctr_dev);
            HIP_TRY(hipEventRecord(c.ev[6], c.stream));
            dispatch(kind, dem, [&](auto K, auto D) {
                launch_march<decltype(K)::value, decltype(D)::value>(grid_march, stream, d, b, M);
            });
            HIP_TRY(hipEventRecord(c.ev[3], c.stream));
            dispatch(kind, dem, [&](auto K, auto D) {
                launch_finish<decltype(K)::value, decltype(D)::value>(grid_flat, stream, d, b, 0);
            });
            HIP_TRY(hipEventRecord(c.ev[7], c.stream));
            dispatch(kind, dem, [&](auto K, auto D) {
                launch_normals<decltype(K)::value, decltype(D)::value>(grid_flat, stream, d, b, fp32_stencil_ok(p),
                                                                       nullptr, 1, ext_modes(p));
            });
        }
        HIP_TRY(hipGetLastError());
    } else {
        hipLaunchKernelGGL(k_collect_hits, dim3(grid_flat), dim3(256), 0, c.stream, d, b);
        HIP_TRY(hipGetLastError());
    }
    HIP_TRY(hipEventRecord(c.ev[1], c.stream));
    if (want_shadow || want_ao) {
        HIP_TRY(hipStreamSynchronize(c.stream));
        if (progress) progress(user, 1, 2);
        if (cancel && cancel(user)) return set_error(MB3D_HIP_CANCELLED, "cancelled");
    }
    if (want_shadow) {
        ShadowJob j;
        std::memset(&j, 0, sizeof j);
        j.n = post->n_lights;
        for (int l = 0; l < j.n; l++) {
            if (post->light_idx[l] < 0 || post->light_idx[l] > 5) return set_error(MB3D_HIP_ERR_INVALID, "bad light index");
            j.bit[l] = 0x400 << post->light_idx[l];
            for (int k = 0; k < 3; k++) j.vec[l][k] = post->light_vec[l][k];
        }
        j.max_len_mul = post->shadow_max_len_mul;
        dispatch(kind, dem, [&](auto K, auto D) {
            if (f64) launch_shadow<decltype(K)::value, decltype(D)::value, false>(grid_flat, stream, d, b, j);
            else launch_shadow<decltype(K)::value, decltype(D)::value, true>(grid_flat, stream, d, b, j);
        });
        HIP_TRY(hipGetLastError());
    }
    if (want_ao) {
        AoRays r;
        build_rays(post->deao_quality, post->deao_max_len, r);
        const int side = post->deao_dither + 1;
        build_ao_sets(post->deao_quality, post->deao_dither, c.ao_sets);
        AoSet *dsets = (AoSet *)((char *)work + ao_off);
        HIP_TRY(hipMemcpyAsync(dsets, c.ao_sets, (size_t)side * side * sizeof(AoSet), hipMemcpyHostToDevice,
                               c.stream));
        r.sets = dsets;
        r.side = side;
        r.rnd = nullptr;
        if (ao_rnd) {
            uint32_t *rnd = (uint32_t *)((char *)work + ao_off + ao_sets_bytes);
            AoSeedJob sj;
            std::memset(&sj, 0, sizeof sj);
            sj.side = side;
            for (int yc = 0; yc < side; yc++) sj.rc_y[yc] = c.ao_sets[yc * side].count;
            hipLaunchKernelGGL(k_ao_seed, dim3(rows), dim3(256), 0, c.stream, d, (const uint32_t *)b.gbuf, rnd, sj);
            HIP_TRY(hipGetLastError());
            r.rnd = rnd;
        }
        DevP d5 = d;
        d5.de_add_steps = 5;  // DeaoRows::new (src/deao.rs:254-260)
        dispatch(kind, dem, [&](auto K, auto D) {
            if (f64) launch_deao<decltype(K)::value, decltype(D)::value, false>(grid_flat, stream, d5, b, r);
            else launch_deao<decltype(K)::value, decltype(D)::value, true>(grid_flat, stream, d5, b, r);
        });
        HIP_TRY(hipGetLastError());
    }
    HIP_TRY(hipEventRecord(c.ev[2], c.stream));
    // results to the caller
    if (!dev_out) {
        if (post_only) {
            HIP_TRY(hipMemcpyAsync(post_gbuf, b.gbuf, npix * 20, hipMemcpyDeviceToHost, c.stream));
        } else if (out) {
            if (out->depth) HIP_TRY(hipMemcpyAsync(out->depth, b.depth, npix * 4, hipMemcpyDeviceToHost, c.stream));
            if (out->depth64) HIP_TRY(hipMemcpyAsync(out->depth64, b.depth64, npix * 8, hipMemcpyDeviceToHost, c.stream));
            if (out->iters) HIP_TRY(hipMemcpyAsync(out->iters, b.iters, npix * 2, hipMemcpyDeviceToHost, c.stream));
            if (out->normals) HIP_TRY(hipMemcpyAsync(out->normals, b.normals, npix * 12, hipMemcpyDeviceToHost, c.stream));
            if (out->gbuf) HIP_TRY(hipMemcpyAsync(out->gbuf, b.gbuf, npix * 20, hipMemcpyDeviceToHost, c.stream));
        }
    }
    unsigned long long ctr[C_WORDS];
    HIP_TRY(hipMemcpyAsync(ctr, ctr_dev, sizeof ctr, hipMemcpyDeviceToHost, c.stream));
    HIP_TRY(hipStreamSynchronize(c.stream));
    float ms_march = 0, ms_post = 0, ms_a = 0, ms_mlp = 0, ms_cone = 0, ms_nrm = 0;
    HIP_TRY(hipEventElapsedTime(&ms_march, c.ev[0], c.ev[1]));
    HIP_TRY(hipEventElapsedTime(&ms_post, c.ev[1], c.ev[2]));
    const bool split = !post_only && !f64;
    if (split) {
        HIP_TRY(hipEventElapsedTime(&ms_a, c.ev[0], c.ev[3]));
        HIP_TRY(hipEventElapsedTime(&ms_nrm, c.ev[7], c.ev[1]));
        float ms_tab = 0;
        HIP_TRY(hipEventElapsedTime(&ms_tab, c.ev[0], c.ev[4]));
        HIP_TRY(hipEventElapsedTime(&ms_cone, c.ev[4], c.ev[5]));
        HIP_TRY(hipEventElapsedTime(&ms_mlp, c.ev[5], c.ev[6]));
        ms_mlp += ms_tab;  // (the ray tables, a few microseconds, with the network)
    }
    mb3d_hip_counters &t = t_counters;
    const double paint_ms = t.paint_ms;  // the last paint's, kept
    std::memset(&t, 0, sizeof t);
    t.paint_ms = paint_ms;
    t.pixels = npix;
    t.steps = ctr[C_STATS + 0];
    t.de_evals_f32 = ctr[C_STATS + 1];
    t.de_evals_f64 = ctr[C_STATS + 2];
    t.its_p8_f32 = ctr[C_STATS + 3];
    t.its_box_f32 = ctr[C_STATS + 4];
    t.its_p8_f64 = ctr[C_STATS + 5];
    t.its_box_f64 = ctr[C_STATS + 6];
    t.hits = ctr[C_STATS + 7];
    t.overflow = ctr[C_STATS + 8];
    t.jumps = ctr[C_STATS + 9];
    t.rollbacks = ctr[C_STATS + 10];
    t.its_other_f64 = ctr[C_STATS + 11];
    t.its_normals_f32 = ctr[C_STATS + 12];
    t.cone_rays = ctr[C_STATS + 13];
    t.cone_background = ctr[C_STATS + 14];
    t.mlp_rays = ctr[C_STATS + 15];
    t.its_cone_f64 = ctr[C_STATS + 16];
    t.de_cone_f64 = ctr[C_STATS + 17];
    t.steps_f64 = ctr[C_STATS + 18];
    t.kernel_ms = ms_march;
    t.post_ms = ms_post;
    t.normals_ms = split ? ms_nrm : 0.0;
    t.mlp_ms = split ? ms_mlp : 0.0;
    t.cone_ms = split ? ms_cone : 0.0;
    t.march_f32_ms = split ? ms_a - ms_mlp - ms_cone : 0.0;
    t.finish_f64_ms = split ? ms_march - ms_a : ms_march;
    if (progress) progress(user, 2, 2);
    // the overflow verdict (src/gpu.rs:2350-2375): the f64 finish already
    // recomputed those pixels, so it is reported, not failed
    return MB3D_HIP_OK;
}

// ---------------------------------------------------------------------------
// batches: many frames (any sizes, the same formula) in one persistent
// launch of each pass, optionally captured into a HIP graph

struct Plan {
    int n = 0;
    int kind = 0, dem = 0;
    int grid_march = 0, grid_flat = 0;
    long long npix = 0;
    DevP P0;
    std::vector<DevP> tab;
    DevP *d_tab = nullptr;
    void *d_work = nullptr;
    unsigned long long *d_ctr = nullptr;
    Bufs b;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    // e0 start, ea after the ray tables and the cone rays, eb after the
    // network, e1 after the FP32 march, ef after the finish, e2 end
    hipEvent_t e0 = nullptr, e1 = nullptr, e2 = nullptr, ea = nullptr, eb = nullptr, ef = nullptr;
    MlpDev M;  // the warm start (M.enabled = 0: none)
    bool n32 = true;  // the normals' stencil in FP32 (every frame allows it)
    bool ext = false;  // a frame asks for the general surface modes (k_normals_ext)
};

void plan_free(Plan *pl) {
    if (!pl) return;
    if (pl->exec) (void)hipGraphExecDestroy(pl->exec);
    if (pl->graph) (void)hipGraphDestroy(pl->graph);
    if (pl->d_tab) (void)hipFree(pl->d_tab);
    if (pl->d_work) (void)hipFree(pl->d_work);
    if (pl->d_ctr) (void)hipFree(pl->d_ctr);
    if (pl->e0) (void)hipEventDestroy(pl->e0);
    if (pl->e1) (void)hipEventDestroy(pl->e1);
    if (pl->e2) (void)hipEventDestroy(pl->e2);
    if (pl->ea) (void)hipEventDestroy(pl->ea);
    if (pl->eb) (void)hipEventDestroy(pl->eb);
    if (pl->ef) (void)hipEventDestroy(pl->ef);
    delete pl;
}

// the kernels of one batch run, in stream order; `split` records the event
// between the two passes (not in a captured graph: an event node there does
// not time)
int plan_enqueue(Plan &pl, hipStream_t stream, bool split) {
    hipLaunchKernelGGL(k_zero, dim3(1), dim3(64), 0, stream, pl.d_ctr, (int)C_WORDS);
    launch_ray_tab(stream, pl.P0, pl.d_tab, pl.n, const_cast<double2 *>(pl.b.rtab));
    const DevP *tab = pl.d_tab;
    const int n = pl.n;
    if (pl.P0.cone_mode) {
        dispatch(pl.kind, pl.dem, [&](auto K, auto D) {
            launch_cone<decltype(K)::value, decltype(D)::value, true>(stream, pl.P0, tab, n, pl.b.rtab, pl.b.cone,
                                                                     pl.d_ctr, pl.M, pl.b.prop);
        });
    }
    if (split) HIP_TRY(hipEventRecord(pl.ea, stream));
    if (jumps_on(pl.M))
        launch_propose<true>(stream, pl.P0, pl.d_tab, pl.n, pl.M, pl.b.rtab, pl.b.cone, pl.b.prop, pl.d_ctr);
    if (split) HIP_TRY(hipEventRecord(pl.eb, stream));
    dispatch(pl.kind, pl.dem, [&](auto K, auto D) {
        launch_march<decltype(K)::value, decltype(D)::value, true>(pl.grid_march, stream, pl.P0, pl.b, pl.M, tab, n);
    });
    if (split) HIP_TRY(hipEventRecord(pl.e1, stream));
    dispatch(pl.kind, pl.dem, [&](auto K, auto D) {
        launch_finish<decltype(K)::value, decltype(D)::value, true>(pl.grid_flat, stream, pl.P0, pl.b, 0, tab, n);
    });
    if (split) HIP_TRY(hipEventRecord(pl.ef, stream));
    dispatch(pl.kind, pl.dem, [&](auto K, auto D) {
        launch_normals<decltype(K)::value, decltype(D)::value, true>(pl.grid_flat, stream, pl.P0, pl.b, pl.n32, tab,
                                                                     n, pl.ext);
    });
    HIP_TRY(hipGetLastError());
    return 0;
}

int plan_build(const mb3d_hip_params *frames, int n, const mb3d_hip_out *out, const mb3d_hip_options *opt_in,
               bool graph, Plan **result) {
    Ctx &c = g_ctx;
    if (c.device < 0) return set_error(MB3D_HIP_ERR_NOT_INIT, "mb3d_hip_init was not called");
    if (!frames || n <= 0 || !out || !result) return set_error(MB3D_HIP_ERR_INVALID, "bad argument");
    if (n > MAX_BATCH_FRAMES) return set_error(MB3D_HIP_ERR_INVALID, "a batch holds at most 64 frames");
    if (!out->on_device) return set_error(MB3D_HIP_ERR_INVALID, "a batch writes device buffers (out->on_device = 1)");
    mb3d_hip_options o;
    std::memset(&o, 0, sizeof o);
    if (opt_in) o = *opt_in;
    if (!tile_ok(o)) return set_error(MB3D_HIP_ERR_INVALID, "bad tile shape (sides are powers of two up to 64)");
    if (!gate_ok(o))
        return set_error(MB3D_HIP_ERR_INVALID, "mlp_cone_gate is a probability in (0, 1) and needs mlp and mlp_cone");
    const int kind = kind_of(frames[0]), dem = dem_of(frames[0]);
    for (int f = 0; f < n; f++) {
        const std::string why = unsupported_reason(&frames[f]);
        if (!why.empty()) return set_error(MB3D_HIP_UNSUPPORTED, why.c_str());
        if (kind_of(frames[f]) != kind || dem_of(frames[f]) != dem)
            return set_error(MB3D_HIP_ERR_INVALID, "the frames of a batch share one formula and DE mode");
        if (o.precision == MB3D_HIP_PREC_F64 || (o.precision == MB3D_HIP_PREC_AUTO && needs_f64(frames[f])) ||
            kind_of(frames[f]) == K_GEN)
            return set_error(MB3D_HIP_UNSUPPORTED, "a batch runs the FP32 march (a frame needs FP64)");
    }
    Plan *pl = new Plan();
    pl->n = n;
    pl->kind = kind;
    for (int f = 0; f < n; f++) {
        pl->n32 = pl->n32 && fp32_stencil_ok(frames[f]);
        pl->ext = pl->ext || ext_modes(frames[f]);
    }
    pl->dem = dem;
    pl->tab.resize(n);
    long long q = 0, pix = 0, rt = 0;
    for (int f = 0; f < n; f++) {
        DevP &d = pl->tab[f];
        fill_devp(frames[f], o, 0, frames[f].height, d);
        d.q_begin = q;
        d.pix_base = pix;
        d.rt_base = (int)rt;
        q += d.nslots_q;
        pix += d.npix;
        rt += ray_tab_entries(d);
    }
    if (pix >= (1ll << 31)) {
        plan_free(pl);
        return set_error(MB3D_HIP_ERR_INVALID, "a batch holds at most 2^31 pixels");
    }
    if (pl->tab[0].cone_mode != 0 && (q >> pl->tab[0].tn_shift) >= (1ll << LIVE_TILE_BITS)) {
        plan_free(pl);
        return set_error(MB3D_HIP_ERR_INVALID, "with the cone rays a batch holds at most 2^26 tiles");
    }
    pl->npix = pix;
    mlp_dev(o.mlp, o, pl->M);
    const int wpc = o.waves_per_cu > 0 ? o.waves_per_cu : 16;
    const size_t cap = list_capacity(pix, c.cus, wpc);
    const size_t base_need = (cap * (sizeof(Cand) + sizeof(SurfRec) + sizeof(int32_t)) + 4096 + 255) & ~(size_t)255;
    pl->P0 = pl->tab[0];
    pl->P0.nslots_q = q;
    pl->P0.npix = pix;
    pl->grid_march = std::max(1, c.cus * wpc / 4);
    pl->grid_flat = std::max(1, c.cus * 4);
    auto fail_free = [&](int e) {
        plan_free(pl);
        return e;
    };
    if (hipMalloc((void **)&pl->d_tab, sizeof(DevP) * n) != hipSuccess ||
        hipMalloc(&pl->d_work, base_need + (pl->M.enabled ? learn_bufs(nullptr, q, nullptr) : 0) +
                                   (size_t)rt * sizeof(double2) +
                                   (pl->P0.cone_mode ? (size_t)(q >> pl->P0.tn_shift) * sizeof(ConeRes) : 0)) !=
            hipSuccess ||
        hipMalloc((void **)&pl->d_ctr, C_WORDS * sizeof(unsigned long long)) != hipSuccess)
        return fail_free(set_error(MB3D_HIP_ERR_DEVICE, "hipMalloc failed for a batch"));
    if (hipMemcpy(pl->d_tab, pl->tab.data(), sizeof(DevP) * n, hipMemcpyHostToDevice) != hipSuccess)
        return fail_free(set_error(MB3D_HIP_ERR_DEVICE, "hipMemcpy failed for a batch"));
    if (hipEventCreate(&pl->e0) != hipSuccess || hipEventCreate(&pl->e1) != hipSuccess ||
        hipEventCreate(&pl->e2) != hipSuccess || hipEventCreate(&pl->ea) != hipSuccess ||
        hipEventCreate(&pl->eb) != hipSuccess || hipEventCreate(&pl->ef) != hipSuccess)
        return fail_free(set_error(MB3D_HIP_ERR_DEVICE, "hipEventCreate failed"));
    std::memset(&pl->b, 0, sizeof pl->b);
    pl->b.depth = out->depth;
    pl->b.depth64 = out->depth64;
    pl->b.iters = out->iters;
    pl->b.normals = out->normals;
    pl->b.gbuf = out->gbuf;
    pl->b.cand = (Cand *)pl->d_work;
    pl->b.surf = (SurfRec *)((char *)pl->d_work + cap * sizeof(Cand));
    pl->b.hits = (int32_t *)(pl->b.surf + cap);
    if (pl->M.enabled) learn_bufs((char *)pl->d_work + base_need, q, &pl->b);
    pl->b.rtab = (const double2 *)((char *)pl->d_work + base_need + (pl->M.enabled ? learn_bufs(nullptr, q, nullptr) : 0));
    pl->b.cone = pl->P0.cone_mode ? (ConeRes *)((char *)pl->b.rtab + (size_t)rt * sizeof(double2)) : nullptr;
    pl->b.ctr = pl->d_ctr;
    if (graph) {
        if (hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal) != hipSuccess)
            return fail_free(set_error(MB3D_HIP_ERR_DEVICE, "hipStreamBeginCapture failed"));
        const int e = plan_enqueue(*pl, c.stream, false);
        hipGraph_t g = nullptr;
        const hipError_t ec = hipStreamEndCapture(c.stream, &g);
        pl->graph = g;
        if (e) return fail_free(e);
        if (ec != hipSuccess) return fail_free(set_error(MB3D_HIP_ERR_DEVICE, "hipStreamEndCapture failed"));
        if (hipGraphInstantiate(&pl->exec, pl->graph, nullptr, nullptr, 0) != hipSuccess)
            return fail_free(set_error(MB3D_HIP_ERR_DEVICE, "hipGraphInstantiate failed"));
    }
    *result = pl;
    return MB3D_HIP_OK;
}

// runs a plan `repeats` times and waits; the counters are the last run's,
// the times the mean of the runs (the pass split only without a graph)
int plan_run(Plan &pl, int repeats) {
    Ctx &c = g_ctx;
    HIP_TRY(hipEventRecord(pl.e0, c.stream));
    for (int r = 0; r < repeats; r++) {
        if (pl.exec) {
            HIP_TRY(hipGraphLaunch(pl.exec, c.stream));
        } else if (int e = plan_enqueue(pl, c.stream, repeats == 1)) {
            return e;
        }
    }
    HIP_TRY(hipEventRecord(pl.e2, c.stream));
    unsigned long long ctr[C_WORDS];
    HIP_TRY(hipMemcpyAsync(ctr, pl.d_ctr, sizeof ctr, hipMemcpyDeviceToHost, c.stream));
    HIP_TRY(hipStreamSynchronize(c.stream));
    float ms_all = 0, ms_a = -1.0f, ms_mlp = -1.0f, ms_cone = -1.0f, ms_nrm = -1.0f;
    if (hipEventElapsedTime(&ms_all, pl.e0, pl.e2) != hipSuccess) ms_all = -1.0f;
    if (!pl.exec && repeats == 1) {
        if (hipEventElapsedTime(&ms_a, pl.e0, pl.e1) != hipSuccess) ms_a = -1.0f;
        if (hipEventElapsedTime(&ms_nrm, pl.ef, pl.e2) != hipSuccess) ms_nrm = -1.0f;
        if (hipEventElapsedTime(&ms_cone, pl.e0, pl.ea) != hipSuccess) ms_cone = -1.0f;
        if (hipEventElapsedTime(&ms_mlp, pl.ea, pl.eb) != hipSuccess) ms_mlp = -1.0f;
    }
    (void)hipGetLastError();
    mb3d_hip_counters &t = t_counters;
    const double paint_ms = t.paint_ms;  // the last paint's, kept
    std::memset(&t, 0, sizeof t);
    t.paint_ms = paint_ms;
    t.pixels = (uint64_t)pl.npix;
    t.steps = ctr[C_STATS + 0];
    t.de_evals_f32 = ctr[C_STATS + 1];
    t.de_evals_f64 = ctr[C_STATS + 2];
    t.its_p8_f32 = ctr[C_STATS + 3];
    t.its_box_f32 = ctr[C_STATS + 4];
    t.its_p8_f64 = ctr[C_STATS + 5];
    t.its_box_f64 = ctr[C_STATS + 6];
    t.hits = ctr[C_STATS + 7];
    t.overflow = ctr[C_STATS + 8];
    t.jumps = ctr[C_STATS + 9];
    t.rollbacks = ctr[C_STATS + 10];
    t.its_other_f64 = ctr[C_STATS + 11];
    t.its_normals_f32 = ctr[C_STATS + 12];
    t.cone_rays = ctr[C_STATS + 13];
    t.cone_background = ctr[C_STATS + 14];
    t.mlp_rays = ctr[C_STATS + 15];
    t.its_cone_f64 = ctr[C_STATS + 16];
    t.de_cone_f64 = ctr[C_STATS + 17];
    t.steps_f64 = ctr[C_STATS + 18];
    t.kernel_ms = ms_all > 0.0f ? ms_all / repeats : -1.0;
    t.normals_ms = ms_nrm >= 0.0f ? ms_nrm : -1.0;
    const bool split = ms_a > 0.0f && ms_mlp >= 0.0f && ms_cone >= 0.0f;
    t.mlp_ms = split ? ms_mlp : -1.0;
    t.cone_ms = split ? ms_cone : -1.0;
    t.march_f32_ms = split ? ms_a - ms_mlp - ms_cone : -1.0;
    t.finish_f64_ms = ms_a > 0.0f ? t.kernel_ms - ms_a : -1.0;
    return MB3D_HIP_OK;
}

// The project's autotuner: every tile shape, persistent grid size,
// iteration-cap split and refill threshold of the sweep on a graph of the
// batch, the fastest kept (min of 3 replays each).
int autotune_locked(const mb3d_hip_params *frames, int n, const mb3d_hip_out *out, mb3d_hip_options *best,
                    double *best_ms) {
    static const int tiles[][2] = {{8, 8}, {16, 4}, {16, 8}, {32, 4}, {32, 8}};
    static const int wpcs[] = {12, 16, 20, 24};
    static const int splits[] = {0, 128, 256};
    static const int refills[] = {8, 16};
    mb3d_hip_options o;
    std::memset(&o, 0, sizeof o);
    if (best) o = *best;
    double top = 1e300;
    mb3d_hip_options win = o;
    for (const auto &t : tiles) {
        for (int w : wpcs) {
            for (int sp : splits)
            for (int rf : refills) {
                mb3d_hip_options c = o;
                c.tile_w = t[0];
                c.tile_h = t[1];
                c.waves_per_cu = w;
                c.it_split = sp;
                c.refill_min = rf;
                Plan *pl = nullptr;
                if (int e = plan_build(frames, n, out, &c, true, &pl)) return e;
                double m = 1e300;
                for (int r = 0; r < 3; r++) {
                    if (int e = plan_run(*pl, 1)) {
                        plan_free(pl);
                        return e;
                    }
                    m = std::min(m, t_counters.kernel_ms);
                }
                plan_free(pl);
                if (m < top) {
                    top = m;
                    win = c;
                }
            }
        }
    }
    if (best) *best = win;
    if (best_ms) *best_ms = top;
    return MB3D_HIP_OK;
}

}  // namespace mb3d

using namespace mb3d;

extern "C" const char *mb3d_hip_last_error(void) { return t_error.c_str(); }

extern "C" int32_t mb3d_hip_abi_version(void) { return MB3D_HIP_ABI_VERSION; }

extern "C" int32_t mb3d_hip_init(int32_t device) {
    try {
        std::lock_guard<std::mutex> lk(g_ctx.mu);
        return init_locked(device);
    } catch (...) {
        return set_error(MB3D_HIP_ERR_DEVICE, "unexpected failure in mb3d_hip_init");
    }
}

extern "C" void mb3d_hip_shutdown(void) {
    std::lock_guard<std::mutex> lk(g_ctx.mu);
    Ctx &c = g_ctx;
    if (c.device < 0) return;
    (void)hipStreamSynchronize(c.stream);
    if (c.d_out) (void)hipFree(c.d_out);
    if (c.d_work) (void)hipFree(c.d_work);
    if (c.d_ctr) (void)hipFree(c.d_ctr);
    for (auto &e : c.ev)
        if (e) (void)hipEventDestroy(e);
    if (c.stream) (void)hipStreamDestroy(c.stream);
    c.d_out = c.d_work = nullptr;
    c.d_ctr = nullptr;
    c.out_bytes = c.work_bytes = 0;
    c.stream = nullptr;
    for (auto &e : c.ev) e = nullptr;
    c.device = -1;
}

extern "C" int32_t mb3d_hip_march(const mb3d_hip_params *params, const mb3d_hip_post_job *post, int32_t row0,
                                  int32_t rows, const mb3d_hip_out *out, const mb3d_hip_options *options,
                                  mb3d_hip_progress_fn progress, mb3d_hip_cancel_fn cancel, void *user) {
    try {
        std::lock_guard<std::mutex> lk(g_ctx.mu);
        return march_locked(params, post, row0, rows, out, options, progress, cancel, user, nullptr, 0);
    } catch (...) {
        return set_error(MB3D_HIP_ERR_DEVICE, "unexpected failure in mb3d_hip_march");
    }
}

extern "C" int32_t mb3d_hip_march_noalloc(const mb3d_hip_params *params, int32_t row0, int32_t rows,
                                          const mb3d_hip_out *out, const mb3d_hip_options *options) {
    try {
        std::lock_guard<std::mutex> lk(g_ctx.mu);
        if (!out) return set_error(MB3D_HIP_ERR_INVALID, "out is NULL");
        return march_locked(params, nullptr, row0, rows, out, options, nullptr, nullptr, nullptr, nullptr, 0, true);
    } catch (...) {
        return set_error(MB3D_HIP_ERR_DEVICE, "unexpected failure in mb3d_hip_march_noalloc");
    }
}

extern "C" int32_t mb3d_hip_post(const mb3d_hip_params *params, const mb3d_hip_post_job *post, int32_t row0,
                                 int32_t rows, uint32_t *gbuf, int32_t on_device, const mb3d_hip_options *options) {
    try {
        if (!gbuf || !post) return set_error(MB3D_HIP_ERR_INVALID, "NULL argument");
        std::lock_guard<std::mutex> lk(g_ctx.mu);
        return march_locked(params, post, row0, rows, nullptr, options, nullptr, nullptr, nullptr, gbuf, on_device);
    } catch (...) {
        return set_error(MB3D_HIP_ERR_DEVICE, "unexpected failure in mb3d_hip_post");
    }
}

extern "C" int32_t mb3d_hip_de_at(const mb3d_hip_params *params, const double *points, int64_t n, double *de,
                                  int32_t *its) {
    try {
        std::lock_guard<std::mutex> lk(g_ctx.mu);
        Ctx &c = g_ctx;
        if (c.device < 0) return set_error(MB3D_HIP_ERR_NOT_INIT, "mb3d_hip_init was not called");
        if (!params || !points || n < 0) return set_error(MB3D_HIP_ERR_INVALID, "bad argument");
        const std::string why = unsupported_reason(params);
        if (!why.empty()) return set_error(MB3D_HIP_UNSUPPORTED, why.c_str());
        if (n == 0) return MB3D_HIP_OK;
        mb3d_hip_options o;
        std::memset(&o, 0, sizeof o);
        DevP d;
        fill_devp(*params, o, 0, params->height, d);
        const size_t need = (size_t)n * (24 + 8 + 4) + 1024;
        if (int e = ensure(&c.d_work, &c.work_bytes, need)) return e;
        char *w = (char *)c.d_work;
        double *dp = (double *)w;
        double *dd = (double *)(w + (size_t)n * 24);
        int32_t *di = (int32_t *)(w + (size_t)n * 32);
        HIP_TRY(hipMemcpyAsync(dp, points, (size_t)n * 24, hipMemcpyHostToDevice, c.stream));
        const int grid = (int)std::min<long long>((n + 255) / 256, (long long)c.cus * 8);
        hipStream_t stream = c.stream;
        dispatch(kind_of(*params), dem_of(*params), [&](auto K, auto) {
            launch_de_at<decltype(K)::value>(grid, stream, d, dp, (long long)n, dd, di);
        });
        HIP_TRY(hipGetLastError());
        if (de) HIP_TRY(hipMemcpyAsync(de, dd, (size_t)n * 8, hipMemcpyDeviceToHost, c.stream));
        if (its) HIP_TRY(hipMemcpyAsync(its, di, (size_t)n * 4, hipMemcpyDeviceToHost, c.stream));
        HIP_TRY(hipStreamSynchronize(c.stream));
        return MB3D_HIP_OK;
    } catch (...) {
        return set_error(MB3D_HIP_ERR_DEVICE, "unexpected failure in mb3d_hip_de_at");
    }
}

extern "C" int32_t mb3d_hip_march_rays(const mb3d_hip_params *params, const mb3d_hip_ray_task *tasks, int64_t n,
                                       int32_t depth_sqrt, const mb3d_hip_ray_opts *opts, mb3d_hip_ray_hit *hits,
                                       int32_t on_device) {
    try {
        std::lock_guard<std::mutex> lk(g_ctx.mu);
        Ctx &c = g_ctx;
        if (c.device < 0) return set_error(MB3D_HIP_ERR_NOT_INIT, "mb3d_hip_init was not called");
        if (!params || !tasks || !hits || n < 0) return set_error(MB3D_HIP_ERR_INVALID, "bad argument");
        const std::string why = rays_unsupported_reason(params);
        if (!why.empty()) return set_error(MB3D_HIP_UNSUPPORTED, why.c_str());
        RayPost post;
        const bool with_post = opts && (opts->n_lights > 0 || opts->ao);
        if (opts)
            if (int e = ray_post_of(*opts, post)) return e;
        if (n == 0) return MB3D_HIP_OK;
        const uint32_t npix = (uint32_t)params->width * (uint32_t)params->height;
        if (!on_device)
            for (int64_t i = 0; i < n; i++)
                if (tasks[i].pixel >= npix) return set_error(MB3D_HIP_ERR_INVALID, "a ray's pixel is outside the picture");
        mb3d_hip_options o;
        std::memset(&o, 0, sizeof o);
        DevP d;
        fill_devp(*params, o, 0, params->height, d);
        const size_t tb = (size_t)n * sizeof(mb3d_hip_ray_task), hb = (size_t)n * sizeof(mb3d_hip_ray_hit);
        const mb3d_hip_ray_task *dt = tasks;
        mb3d_hip_ray_hit *dh = hits;
        if (!on_device) {
            const size_t off = (tb + 255) & ~(size_t)255;
            if (int e = ensure(&c.d_work, &c.work_bytes, off + hb + 256)) return e;
            dt = (const mb3d_hip_ray_task *)c.d_work;
            dh = (mb3d_hip_ray_hit *)((char *)c.d_work + off);
            HIP_TRY(hipMemcpyAsync((void *)dt, tasks, tb, hipMemcpyHostToDevice, c.stream));
        }
        hipStream_t stream = c.stream;
        hipLaunchKernelGGL(k_zero, dim3(1), dim3(64), 0, stream, c.d_ctr, (int)C_WORDS);
        HIP_TRY(hipEventRecord(c.ev[0], stream));
        const int grid = (int)std::min<long long>((n + 255) / 256, (long long)c.cus * 8);
        unsigned long long *ctr = c.d_ctr;
        dispatch(kind_of(*params), dem_of(*params), [&](auto K, auto D) {
            launch_rays<decltype(K)::value, decltype(D)::value>(grid, stream, d, dt, (long long)n, depth_sqrt, dh, ctr);
            if (with_post)
                launch_ray_post<decltype(K)::value, decltype(D)::value>(grid, stream, d, dt, (long long)n, post, dh,
                                                                        ctr);
        });
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipEventRecord(c.ev[1], stream));
        if (!on_device) HIP_TRY(hipMemcpyAsync(hits, dh, hb, hipMemcpyDeviceToHost, stream));
        unsigned long long cw[C_WORDS];
        HIP_TRY(hipMemcpyAsync(cw, c.d_ctr, sizeof cw, hipMemcpyDeviceToHost, stream));
        HIP_TRY(hipStreamSynchronize(stream));
        float ms = 0.0f;
        HIP_TRY(hipEventElapsedTime(&ms, c.ev[0], c.ev[1]));
        mb3d_hip_counters &t = t_counters;
        const double paint_ms = t.paint_ms;
        std::memset(&t, 0, sizeof t);
        t.paint_ms = paint_ms;
        t.pixels = (uint64_t)n;
        t.steps = cw[C_STATS + 0];
        t.de_evals_f64 = cw[C_STATS + 2];
        t.its_p8_f64 = cw[C_STATS + 5];
        t.its_box_f64 = cw[C_STATS + 6];
        t.hits = cw[C_STATS + 7];
        t.its_other_f64 = cw[C_STATS + 11];
        t.kernel_ms = ms;
        t.finish_f64_ms = ms;
        return MB3D_HIP_OK;
    } catch (...) {
        return set_error(MB3D_HIP_ERR_DEVICE, "unexpected failure in mb3d_hip_march_rays");
    }
}

extern "C" int32_t mb3d_hip_mc_batch(const mb3d_hip_params *params, const mb3d_hip_mc_task *tasks, int64_t n,
                                     int32_t secant, mb3d_hip_mc_res *res, int32_t on_device) {
    try {
        std::lock_guard<std::mutex> lk(g_ctx.mu);
        Ctx &c = g_ctx;
        if (c.device < 0) return set_error(MB3D_HIP_ERR_NOT_INIT, "mb3d_hip_init was not called");
        if (!params || !tasks || !res || n < 0) return set_error(MB3D_HIP_ERR_INVALID, "bad argument");
        const std::string why = rays_unsupported_reason(params);
        if (!why.empty()) return set_error(MB3D_HIP_UNSUPPORTED, why.c_str());
        if (n == 0) return MB3D_HIP_OK;
        if (!on_device)
            for (int64_t i = 0; i < n; i++)
                if (tasks[i].kind < MB3D_HIP_MC_RAY || tasks[i].kind > MB3D_HIP_MC_THROUGH)
                    return set_error(MB3D_HIP_ERR_INVALID, "a task's kind is not 0..2");
        mb3d_hip_options o;
        std::memset(&o, 0, sizeof o);
        DevP d;
        fill_devp(*params, o, 0, params->height, d);
        const size_t tb = (size_t)n * sizeof(mb3d_hip_mc_task), rb = (size_t)n * sizeof(mb3d_hip_mc_res);
        const mb3d_hip_mc_task *dt = tasks;
        mb3d_hip_mc_res *dr = res;
        if (!on_device) {
            const size_t off = (tb + 255) & ~(size_t)255;
            if (int e = ensure(&c.d_work, &c.work_bytes, off + rb + 256)) return e;
            dt = (const mb3d_hip_mc_task *)c.d_work;
            dr = (mb3d_hip_mc_res *)((char *)c.d_work + off);
            HIP_TRY(hipMemcpyAsync((void *)dt, tasks, tb, hipMemcpyHostToDevice, c.stream));
        }
        hipStream_t stream = c.stream;
        hipLaunchKernelGGL(k_zero, dim3(1), dim3(64), 0, stream, c.d_ctr, (int)C_WORDS);
        HIP_TRY(hipEventRecord(c.ev[0], stream));
        const int grid = (int)std::min<long long>((n + 255) / 256, (long long)c.cus * 8);
        unsigned long long *ctr = c.d_ctr;
        dispatch(kind_of(*params), dem_of(*params), [&](auto K, auto D) {
            launch_mc<decltype(K)::value, decltype(D)::value>(grid, stream, d, dt, (long long)n, secant != 0, dr, ctr);
        });
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipEventRecord(c.ev[1], stream));
        if (!on_device) HIP_TRY(hipMemcpyAsync(res, dr, rb, hipMemcpyDeviceToHost, stream));
        unsigned long long cw[C_WORDS];
        HIP_TRY(hipMemcpyAsync(cw, c.d_ctr, sizeof cw, hipMemcpyDeviceToHost, stream));
        HIP_TRY(hipStreamSynchronize(stream));
        float ms = 0.0f;
        HIP_TRY(hipEventElapsedTime(&ms, c.ev[0], c.ev[1]));
        mb3d_hip_counters &t = t_counters;
        const double paint_ms = t.paint_ms;
        std::memset(&t, 0, sizeof t);
        t.paint_ms = paint_ms;
        t.pixels = (uint64_t)n;
        t.steps = cw[C_STATS + 0];
        t.de_evals_f64 = cw[C_STATS + 2];
        t.its_p8_f64 = cw[C_STATS + 5];
        t.its_box_f64 = cw[C_STATS + 6];
        t.hits = cw[C_STATS + 7];
        t.its_other_f64 = cw[C_STATS + 11];
        t.kernel_ms = ms;
        t.finish_f64_ms = ms;
        return MB3D_HIP_OK;
    } catch (...) {
        return set_error(MB3D_HIP_ERR_DEVICE, "unexpected failure in mb3d_hip_mc_batch");
    }
}

extern "C" int32_t mb3d_hip_rays(const mb3d_hip_params *params, int32_t row0, int32_t rows, float *rays,
                                 float *step_len) {
    try {
        std::lock_guard<std::mutex> lk(g_ctx.mu);
        Ctx &c = g_ctx;
        if (c.device < 0) return set_error(MB3D_HIP_ERR_NOT_INIT, "mb3d_hip_init was not called");
        if (!params || row0 < 0 || rows < 0 || row0 + rows > params->height)
            return set_error(MB3D_HIP_ERR_INVALID, "bad argument");
        const long long n = (long long)rows * params->width;
        if (n == 0) return MB3D_HIP_OK;
        mb3d_hip_options o;
        std::memset(&o, 0, sizeof o);
        DevP d;
        fill_devp(*params, o, row0, rows, d);
        if (int e = ensure(&c.d_work, &c.work_bytes, (size_t)n * 28 + 1024)) return e;
        float *dr = (float *)c.d_work;
        float *dl = dr + (size_t)n * 6;
        const int grid = (int)std::min<long long>((n + 255) / 256, (long long)c.cus * 8);
        hipLaunchKernelGGL(k_rays, dim3(grid), dim3(256), 0, c.stream, d, dr, dl);
        HIP_TRY(hipGetLastError());
        if (rays) HIP_TRY(hipMemcpyAsync(rays, dr, (size_t)n * 24, hipMemcpyDeviceToHost, c.stream));
        if (step_len) HIP_TRY(hipMemcpyAsync(step_len, dl, (size_t)n * 4, hipMemcpyDeviceToHost, c.stream));
        HIP_TRY(hipStreamSynchronize(c.stream));
        return MB3D_HIP_OK;
    } catch (...) {
        return set_error(MB3D_HIP_ERR_DEVICE, "unexpected failure in mb3d_hip_rays");
    }
}

struct mb3d_hip_plan {
    mb3d::Plan *p;
};

extern "C" int32_t mb3d_hip_march_batch(const mb3d_hip_params *frames, int32_t n, const mb3d_hip_out *out,
                                        const mb3d_hip_options *options) {
    try {
        std::lock_guard<std::mutex> lk(g_ctx.mu);
        Plan *pl = nullptr;
        if (int e = plan_build(frames, n, out, options, false, &pl)) return e;
        const int e = plan_run(*pl, 1);
        plan_free(pl);
        return e;
    } catch (...) {
        return set_error(MB3D_HIP_ERR_DEVICE, "unexpected failure in mb3d_hip_march_batch");
    }
}

extern "C" int32_t mb3d_hip_plan_create(const mb3d_hip_params *frames, int32_t n, const mb3d_hip_out *out,
                                        const mb3d_hip_options *options, mb3d_hip_plan **plan) {
    try {
        if (!plan) return set_error(MB3D_HIP_ERR_INVALID, "plan is NULL");
        *plan = nullptr;
        std::lock_guard<std::mutex> lk(g_ctx.mu);
        Plan *pl = nullptr;
        if (int e = plan_build(frames, n, out, options, true, &pl)) return e;
        *plan = new mb3d_hip_plan{pl};
        return MB3D_HIP_OK;
    } catch (...) {
        return set_error(MB3D_HIP_ERR_DEVICE, "unexpected failure in mb3d_hip_plan_create");
    }
}

extern "C" int32_t mb3d_hip_plan_launch(mb3d_hip_plan *plan, int32_t repeats) {
    try {
        if (!plan || !plan->p || repeats < 1) return set_error(MB3D_HIP_ERR_INVALID, "bad plan or repeat count");
        std::lock_guard<std::mutex> lk(g_ctx.mu);
        return plan_run(*plan->p, repeats);
    } catch (...) {
        return set_error(MB3D_HIP_ERR_DEVICE, "unexpected failure in mb3d_hip_plan_launch");
    }
}

extern "C" void mb3d_hip_plan_destroy(mb3d_hip_plan *plan) {
    if (!plan) return;
    std::lock_guard<std::mutex> lk(g_ctx.mu);
    plan_free(plan->p);
    delete plan;
}

extern "C" int32_t mb3d_hip_autotune(const mb3d_hip_params *frames, int32_t n, const mb3d_hip_out *out,
                                     mb3d_hip_options *best, double *best_ms) {
    try {
        std::lock_guard<std::mutex> lk(g_ctx.mu);
        return autotune_locked(frames, n, out, best, best_ms);
    } catch (...) {
        return set_error(MB3D_HIP_ERR_DEVICE, "unexpected failure in mb3d_hip_autotune");
    }
}

extern "C" int32_t mb3d_hip_counters_get(mb3d_hip_counters *out) {
    if (!out) return set_error(MB3D_HIP_ERR_INVALID, "NULL argument");
    *out = t_counters;
    return MB3D_HIP_OK;
}
