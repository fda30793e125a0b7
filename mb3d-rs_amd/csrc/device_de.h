// device_de.h -- device code of the DE ray marcher for gfx950 (CDNA4, wave64).
//
// Two precisions of the same reference algorithm:
//  * FP32 "bulk" iterate (de32_*): the formula iteration in single precision
//    with the ray position kept in f64 and rounded per evaluation.  The
//    numeric gradient DE (src/calc.rs:1515-1545) runs its 1 + 3 iteration
//    loops fused in one loop over four points, so the three offset points
//    run in lock step with the first (the reference stops them at the first
//    one's iteration count: they cannot run longer).
//  * FP64 (struct M64): the reference's f64 marcher (Marcher in
//    src/calc.rs, Iteration in src/iteration.rs) restated for the device,
//    compiled without contraction so that it follows the reference's
//    operation order.  It makes every decision that sets the output: the
//    stop near the surface, the bisection, the normals, the iteration count.
//
// Formula code: int_pow power 8 (src/formulas.rs:837-858) and amazing_box
// (src/formulas.rs:707-736) in both precisions; every other built-in
// formula (Formula::iterate, src/formulas.rs:576-858), the 4D hybrids and
// Julia mode in FP64 (kind K_GEN).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/mb3d_hip.h"

namespace mb3d {

// K_P8, K_BOX, K_HYB: power-8 Integer Power, Amazing Box and their
// hybrids, with an FP32 iterate; K_GEN: any built-in formula, the 4D
// hybrids and Julia mode, FP64 only.
enum Kind : int { K_P8 = 0, K_BOX = 1, K_HYB = 2, K_GEN = 3 };
enum DeMode : int { DEM_NUMERIC = 0, DEM_BOX = 1, DEM_P8 = 2 };

// The march parameters as the kernels take them (by value, so that the
// uniform ones live in scalar registers).  The fields up to `opt2` are the
// core every kind reads; a batch stages only those in LDS (FrameTab), so the
// fields from `opt2` on belong to the general kind (K_GEN) alone, which
// never runs batched.
struct DevP {
    int width, height, row0, rows;
    int n_slots, end_to, repeat_from, max_it, min_it;
    int de_option, is_custom_de, de_add_steps, dfog_on_it, first_step_random, step_cap;
    int tile_w, tile_h, tiles_x;
    int slot_kind[MB3D_HIP_MAX_SLOTS], slot_iters[MB3D_HIP_MAX_SLOTS];
    // the first three f64 options of each slot, Formula::options order
    // without the integers (src/formulas.rs:308-357), the rest in opt2:
    // Integer Power (z_mul); Amazing Box (scale, min_r, fold); Real Power
    // (power, z_mul); Quaternion (yw_mul, w_add); Tricorn (z_mul, cz_mul);
    // Bulbox (scale, min_r, fold | bulb_scale, r_threshold, r_threshold2);
    // Folding Int Pow (z_mul, fold); Aexion C (power, z_mul, power_c | cz_mul)
    double opt[MB3D_HIP_MAX_SLOTS][3];
    float fz_mul[MB3D_HIP_MAX_SLOTS], fscale[MB3D_HIP_MAX_SLOTS], fmin_r2[MB3D_HIP_MAX_SLOTS];
    float fscale_min_r2[MB3D_HIP_MAX_SLOTS], ffold[MB3D_HIP_MAX_SLOTS];
    float d_de_scale, de_stop, z_step_div, mh04zsd, de_stop_factor, ln_rstop, de_offset, de_offset006;
    float mcts_m, ms_de_sub, d_col_plus, col_var, fovx_off, fovx_mul, guard;
    float f_d_rstop, f_rstop3d;
    float fhln[MB3D_HIP_MAX_SLOTS];
    double d_rstop, rstop3d, fov_y, zcorr, zc_mul, zend, step_width;
    double vg[3][3], ystart[3];
    int64_t npix;      // pixels of the rectangle
    int64_t nslots_q;  // work items of the tile queue (tiles x tile_w x tile_h)
    // a frame of a batch: its first work item in the batch's queue and its
    // first pixel in the batch's output arrays (0 for a single picture)
    int64_t q_begin, pix_base;
    int it_split;  // march steps after which the FP32 pass hands a ray to the f64 finish (0: none)
    // the frame's entries in the ray tables (RayTab): `width` columns, then
    // `rows` rows
    int rt_base;
    // log2 of the tile width and of the tile's item count (tile sides are
    // powers of two)
    int tw_shift, tn_shift;
    // idle lanes a wave of the FP32 march gathers before it refills them
    // (1: every turn that has an idle lane)
    int refill_min;
    // the cone rays (k_cone): 0 off, 1 a tile whose cone ray passes the end
    // of the view is background, 2 also start the tile's rays at its depth
    int cone_mode;
    // the surface's normal and colour modes (surface_found,
    // src/calc.rs:2471-2517): smooth normals 0..8, normals on the DE or on
    // the smooth iteration count, the colour option and color_on_it
    int sm_normals, normals_on_de, color_option, color_on_it;
    float mct_color_mul;
    int pad_core2;

    // ---- the general kind (K_GEN) only ----
    double opt2[MB3D_HIP_MAX_SLOTS][3];
    // Integer Power, Folding Int Pow: the power; Aexion C: rotate C (1),
    // conditional phi (2), power C on |Z - C| (4), mode << 8
    int slot_iarg[MB3D_HIP_MAX_SLOTS];
    int slot_ade[MB3D_HIP_MAX_SLOTS];  // the analytic-DE variant (each part's own flag)
    int slot_unc[MB3D_HIP_MAX_SLOTS];  // iterated but not counted (nHybrid < 0)
    int four_d, do_julia, ipol;
    float ipol_w[2];  // the interpolation hybrid's weights
    double ju[4];     // the Julia constant (and J4)
    float sm4[12];    // rows of smatrix4's first three columns (f32 values)
    // the DE combination (0: none; 1 min, 2 max, 3 max of the inverted
    // second part, 4 and 5 smooth minima) and its second hybrid part
    int decomb, start2, end2, repeat2, max_it2, de_option2, custom2, four_d2, placed;
    float d_de_scale2;
    double smooth;  // the smooth combinations' distance (steps)
    // the second part's space (Place, src/calc.rs:238-253): m (p - offset),
    // its distances times place_scale
    double place_off[3], place_m[3][3], place_scale;
};

// Per-lane work counters, reduced per wave at the end of a kernel.
struct Cnt {
    uint32_t steps = 0, de32 = 0, de64 = 0;
    uint32_t p8_32 = 0, box_32 = 0, p8_64 = 0, box_64 = 0, oth_64 = 0;
    uint32_t nrm_32 = 0;  // of p8_32 and box_32: the normals pass's FP32 stencil
    uint32_t hits = 0, overflow = 0;
    uint32_t jumps = 0, rollbacks = 0;  // warm starts kept, warm starts undone
    uint32_t cone_rays = 0, cone_bg = 0;  // cone rays marched; pixels their cone made background
    uint32_t cone_its = 0, cone_des = 0;  // of the f64 iterations and DEs: the cone rays'
    uint32_t steps64 = 0;                 // of the steps: the f64 finish's
};

// ---------------------------------------------------------------------------
// FP32 formula iterations
//
// The FMAs are written out (fmaf) rather than left to contraction: which
// products the compiler would fuse depends on the code around an inlined
// call, so two kernels (a batch's and a single frame's) could round the same
// DE differently; written out, every kernel computes the same bits.

// The power-8 triplex step (the HybridIntP8 polynomial, src/formulas.rs:
// 837-858) in FP32, arranged for fewer operations with the same values up to
// rounding.  With X = x^2, Y = y^2, R = X + Y, Z = z^2:
//   x^8 - 28 x^6 y^2 + 70 x^4 y^4 - 28 x^2 y^6 + y^8 = R^4 - 32 R^2 XY + 128 (XY)^2
//   x^6 - 7 x^4 y^2 + 7 x^2 y^4 - y^6 = (X - Y)(R^2 - 8 XY)
//   and the polar factor a = 1 + num / R^4 = (r2^4 - 32 r2^2 ZR + 128 (ZR)^2) / R^4
//   (the same Chebyshev identity in R and Z, r2 = R + Z = |v|^2),
// so r2, the squared radius the caller already holds, replaces the fourth
// powers.  (Checked against the f64 reference form on random points: the
// same error distribution as that form evaluated in FP32.)
__device__ __forceinline__ void p8f(float &x, float &y, float &z, float jx, float jy, float jz, float zmul8,
                                    float r2) {
    const float xx = x * x, yy = y * y, zz = z * z;
    const float R = fmaf(x, x, yy), rr = R * R, zr = zz * R;
    // (Z^2 - 6 ZR + R^2) (Z - R) sqrt(R)
    const float t = fmaf(-6.0f, zr, fmaf(zz, zz, rr)) * (zz - R) * __builtin_amdgcn_sqrtf(R);
    const float nz = fmaf(-(t * z), zmul8, jz);
    const float s2 = r2 * r2;
    const float a = fmaf(s2, s2, zr * fmaf(128.0f, zr, -32.0f * s2)) * __builtin_amdgcn_rcpf(fmaf(rr, rr, 1e-37f));
    const float p2 = xx * yy;
    const float ny = fmaf((xx - yy) * fmaf(-8.0f, p2, rr) * (x * y), 8.0f * a, jy);
    const float nx = fmaf(a, fmaf(rr, rr, p2 * fmaf(128.0f, p2, -32.0f * rr)), jx);
    x = nx;
    y = ny;
    z = nz;
}

__device__ __forceinline__ float bfoldf(float x, float f) { return fabsf(x + f) - (fabsf(x - f) + x); }

// x^2 + y^2 + z^2 with two FMAs
__device__ __forceinline__ float sq3f(float x, float y, float z) { return fmaf(z, z, fmaf(y, y, x * x)); }

// amazing_box in FP32; returns the derivative factor
__device__ __forceinline__ float boxf(float &x, float &y, float &z, float jx, float jy, float jz, float fold,
                                     float scale, float min_r2, float scale_min_r2) {
    const float a = bfoldf(x, fold), b = bfoldf(y, fold), c = bfoldf(z, fold);
    const float r = sq3f(c, b, a);
    const float mul = r < min_r2 ? scale_min_r2 : (1.0f < r ? scale : scale * __builtin_amdgcn_rcpf(r));
    x = fmaf(a, mul, jx);
    y = fmaf(b, mul, jy);
    z = fmaf(c, mul, jz);
    return mul;
}

// The slot order of the alternating hybrid (Iteration::run_with,
// src/iteration.rs:276-290): advances (n, btmp) to the slot of the next
// counted iteration.  The sequence is the same for every lane.
__device__ __forceinline__ void next_slot(const DevP &P, int &n, int &btmp) {
    int guard = 0;
    while (btmp <= 0 && guard < 64) {
        n += 1;
        if (n > P.end_to) n = P.repeat_from;
        btmp = P.slot_iters[n];
        guard++;
    }
    btmp -= 1;
}

// ln x in FP32 from the hardware's base-2 logarithm (v_log_f32, ~1 ulp for
// normal x): the FP32 march's DEs steer its steps only, the f64 finish makes
// every decision that reaches an output, so the library logf's scaling for
// denormal inputs and its extended-precision product with ln 2 (a dozen
// instructions a DE) buy nothing here.
__device__ __forceinline__ float ln_fast(float x) { return __builtin_amdgcn_logf(x) * 0.693147180559945309f; }

// The numeric gradient DE in FP32 (calc_de_part, src/calc.rs:1515-1545):
// the point and its three offset points iterated in one loop.
template <int KIND>
__device__ float de32_numeric(const DevP &P, const double c[3], float ms, int &it_out, Cnt &cnt) {
    const double off = (double)P.de_offset;
    float x[4], y[4], z[4], jx[4], jy[4], jz[4];
    jx[0] = (float)c[0];
    jy[0] = (float)c[1];
    jz[0] = (float)c[2];
    jx[1] = (float)(c[0] + off);
    jy[1] = jy[0];
    jz[1] = jz[0];
    jx[2] = jx[0];
    jy[2] = (float)(c[1] + off);
    jz[2] = jz[0];
    jx[3] = jx[0];
    jy[3] = jy[0];
    jz[3] = (float)(c[2] + off);
#pragma unroll
    for (int k = 0; k < 4; k++) {
        x[k] = jx[k];
        y[k] = jy[k];
        z[k] = jz[k];
    }
    float r0 = 0.f, r1 = 0.f, r2 = 0.f, r3 = 0.f;
    float cur[4];  // each point's squared radius now (p8f's r2)
#pragma unroll
    for (int k = 0; k < 4; k++) cur[k] = sq3f(x[k], y[k], z[k]);
    bool a1 = true, a2 = true, a3 = true;
    int it = 0;
    int n = 0, btmp = P.slot_iters[0];
    const float zm8 = 8.0f * P.fz_mul[0];
    for (;;) {
        int kind = KIND;
        int slot = 0;
        if (KIND == K_HYB) {
            next_slot(P, n, btmp);
            slot = n;
            kind = P.slot_kind[n] == MB3D_HIP_FORMULA_AMAZINGBOX ? K_BOX : K_P8;
        }
        if (kind == K_P8) {
            const float zm = KIND == K_HYB ? 8.0f * P.fz_mul[slot] : zm8;
#pragma unroll
            for (int k = 0; k < 4; k++) p8f(x[k], y[k], z[k], jx[k], jy[k], jz[k], zm, cur[k]);
            cnt.p8_32 += 1u + a1 + a2 + a3;
        } else {
            const float fo = P.ffold[slot], sc = P.fscale[slot], m2 = P.fmin_r2[slot], sm = P.fscale_min_r2[slot];
#pragma unroll
            for (int k = 0; k < 4; k++) boxf(x[k], y[k], z[k], jx[k], jy[k], jz[k], fo, sc, m2, sm);
            cnt.box_32 += 1u + a1 + a2 + a3;
        }
        it++;
        r0 = sq3f(x[0], y[0], z[0]);
        const float q1 = sq3f(x[1], y[1], z[1]);
        const float q2 = sq3f(x[2], y[2], z[2]);
        const float q3 = sq3f(x[3], y[3], z[3]);
        cur[0] = r0;
        cur[1] = q1;
        cur[2] = q2;
        cur[3] = q3;
        if (a1) {
            r1 = q1;
            a1 = !(q1 > P.f_rstop3d);
        }
        if (a2) {
            r2 = q2;
            a2 = !(q2 > P.f_rstop3d);
        }
        if (a3) {
            r3 = q3;
            a3 = !(q3 > P.f_rstop3d);
        }
        if (it >= P.max_it || r0 > P.f_d_rstop) break;
    }
    it_out = it;
    cnt.de32++;
    float res;
    if (r0 < 1e-37f) {
        res = 0.0f;
    } else {
        const float g = sq3f(r0 - r1, r0 - r2, r0 - r3);
        res = r0 * ln_fast(r0) * P.d_de_scale * __builtin_amdgcn_rcpf(__builtin_amdgcn_sqrtf(g) + P.de_offset006);
    }
    const float mn = ms * 0.25f;
    if (!(res >= mn)) res = mn;  // also NaN
    return res;
}

// The analytic DEs in FP32: Amazing Box |z| / |dr| (option 11,
// src/iteration.rs:562-568) and the power-8 extension 0.5 r ln r / dr.
template <int DEM>
__device__ float de32_analytic(const DevP &P, const double c[3], float ms, int &it_out, Cnt &cnt,
                               float *otrap = nullptr) {
    const float jx = (float)c[0], jy = (float)c[1], jz = (float)c[2];
    float x = jx, y = jy, z = jz;
    float r = sq3f(x, y, z);
    float ot = r;
    float w = 1.0f;
    int it = 0;
    if (DEM == DEM_BOX) {
        const float fo = P.ffold[0], sc = P.fscale[0], m2 = P.fmin_r2[0], sm = P.fscale_min_r2[0];
        for (;;) {
            w *= boxf(x, y, z, jx, jy, jz, fo, sc, m2, sm);
            it++;
            r = sq3f(x, y, z);
            if (otrap) ot = fminf(ot, r);
            if (it >= P.max_it || r > P.f_d_rstop) break;
        }
        cnt.box_32 += it;
    } else {
        const float zm8 = 8.0f * P.fz_mul[0];
        for (;;) {
            // (r: the squared radius of (x, y, z), from the line before the
            // loop or the end of the last iteration)
            const float r2 = r;
            w = fmaf(8.0f * (r2 * r2 * r2 * __builtin_amdgcn_sqrtf(r2)), w, 1.0f);
            p8f(x, y, z, jx, jy, jz, zm8, r2);
            it++;
            r = sq3f(x, y, z);
            if (otrap) ot = fminf(ot, r);
            if (it >= P.max_it || r > P.f_d_rstop) break;
        }
        cnt.p8_32 += it;
    }
    if (otrap) *otrap = ot;
    it_out = it;
    cnt.de32++;
    float res;
    if (DEM == DEM_BOX)
        res = __builtin_amdgcn_sqrtf(r) * __builtin_amdgcn_rcpf(fabsf(w)) * P.d_de_scale;
    else
        res = __builtin_amdgcn_sqrtf(r) * 0.25f * ln_fast(r) * __builtin_amdgcn_rcpf(w) * P.d_de_scale;
    const float mn = ms * 0.25f;
    if (!(res >= mn)) res = mn;
    return res;
}

template <int KIND, int DEM>
__device__ __forceinline__ float de32(const DevP &P, const double c[3], float ms, int &it, Cnt &cnt) {
    if (DEM == DEM_NUMERIC) return de32_numeric<KIND>(P, c, ms, it, cnt);
    return de32_analytic<DEM>(P, c, ms, it, cnt);
}

// ---------------------------------------------------------------------------
// FP64: the reference's marcher, without contraction

#pragma clang fp contract(off)

__device__ __forceinline__ double bfold64(double x, double f) { return fabs(x + f) - (fabs(x - f) + x); }

// round half to even (f64::round_ties_even)
__device__ __forceinline__ double rte(double x) { return __builtin_rint(x); }
__device__ __forceinline__ float rtef(float x) { return __builtin_rintf(x); }

// min_max_clip_15bit (src/math.rs:246-255)
__device__ __forceinline__ uint32_t clip15(float s) {
    if (isnan(s) || s < 0.0f) return 0;
    if (s > 32767.0f) return 32767;
    return (uint32_t)rtef(s);
}

// The f64 iteration record of one evaluation (Iteration,
// src/iteration.rs:59-105), the parts the built-in formulas use.
struct It64 {
    double v[4], j[4];
    double rout, rold, otrap, deriv1;
    int it_result;
    uint32_t p8, box, oth;
};

// int_pow (src/formulas.rs:769-858), powers 2..7 (8: int_pow8_64)
__device__ __forceinline__ void int_pow_low64(int power, double zmul, double v[4], const double j[4]) {
    const double x = v[0], y = v[1], z = v[2];
    const double EPS = 1e-40;
    const double sx = x * x, sy = y * y, sz = z * z;
    if (power <= 2) {
        const double r = sx + sy;
        v[2] = (sqrt(r) * zmul * z) * 2.0 + j[2];
        const double a = (r - z * z) / r;
        v[0] = (sx - sy) * a + j[0];
        v[1] = (x * y * a) * 2.0 + j[1];
    } else if (power == 3) {
        const double r = sx + sy;
        const double a = 1.0 - 3.0 * sz / (r + EPS);
        v[0] = (sx - 3.0 * sy) * a * x + j[0];
        v[2] = j[2] - (sz - 3.0 * r) * z * zmul;
        v[1] = a * (3.0 * sx - sy) * y + j[1];
    } else if (power == 4) {
        const double r = sx + sy;
        const double a = 1.0 + sz * (sz - 6.0 * r) / (r * r + EPS);
        v[0] = (sx * (sx - 6.0 * sy) + sy * sy) * a + j[0];
        v[2] = sqrt(r) * (r - sz) * z * 4.0 * zmul + j[2];
        v[1] = a * x * (sx - sy) * 4.0 * y + j[1];
    } else if (power == 5) {
        const double r = sx + sy;
        const double a = 1.0 + (sz * sz - sz * r * 2.0) * 5.0 / (r * r + EPS);
        v[1] = (5.0 * sx * sx - sy * (10.0 * sx - sy)) * a * y + j[1];
        v[2] = (sz * (sz - 10.0 * r) + 5.0 * r * r) * z * zmul + j[2];
        v[0] = a * x * (sx * (sx - 10.0 * sy) + 5.0 * sy * sy) + j[0];
    } else if (power == 6) {
        const double r = sx + sy, rr = r * r;
        const double a = 1.0 - sz * (15.0 * rr + sz * (sz - r * 15.0)) / (rr * r + EPS);
        v[1] = (3.0 * sy * sy + sx * (3.0 * sx - 10.0 * sy)) * a * y * x * 2.0 + j[1];
        v[2] = (sz * (3.0 * sz - 10.0 * r) + 3.0 * rr) * sqrt(r) * z * zmul * 2.0 + j[2];
        v[0] = (sy * sy * (15.0 * sx - sy) + sx * sx * (sx - 15.0 * sy)) * a + j[0];
    } else {
        const double r = sx + sy, rr = r * r;
        const double a = 1.0 - 7.0 * (sz * (3.0 * rr + sz * (sz - 5.0 * r))) / (rr * r + EPS);
        v[1] = (sx * (21.0 * sy * sy + sx * (7.0 * sx - 35.0 * sy)) - sy * sy * sy) * a * y + j[1];
        v[2] = j[2] - (sz * sz * sz - 7.0 * r * (rr + sz * (3.0 * sz - 5.0 * r))) * z * zmul;
        v[0] = a * x * (sx * (35.0 * sy * sy + sx * (sx - 21.0 * sy)) - 7.0 * sy * sy * sy) + j[0];
    }
}

// HybridIntP8 (src/formulas.rs:837-858)
__device__ __forceinline__ void int_pow8_64(double zmul, double v[4], const double j[4]) {
    const double x = v[0], y = v[1], z = v[2];
    const double xx = x * x, yy = y * y, zz = z * z;
    const double r = xx + yy;
    const double rr = r * r;
    const double zzzz = zz * zz;
    const double t = (zzzz - 6.0 * r * zz + rr) * (zz - r) * sqrt(r);
    v[2] = -(t * z * 8.0 * zmul) + j[2];
    const double num = (rr * 70.0 + zzzz) * zzzz - 28.0 * zz * r * (zzzz + rr);
    const double a = num / (rr * rr + 1e-40) + 1.0;
    const double xxxx = xx * xx, yyyy = yy * yy;
    v[1] = (yyyy * (7.0 * xx - yy) + xxxx * (xx - 7.0 * yy)) * 8.0 * x * y * a + j[1];
    v[0] = a * (xxxx * xxxx + yyyy * (70.0 * xxxx + yyyy) - 28.0 * (xx * yy * (yyyy + xxxx))) + j[0];
}

__device__ __forceinline__ void int_pow64(int power, double zmul, double v[4], const double j[4]) {
    if (power >= 8)
        int_pow8_64(zmul, v, j);
    else
        int_pow_low64(power, zmul, v, j);
}

// The f64 square root and reciprocal from the hardware's estimates
// (v_rsq_f64, v_rcp_f64) refined by Newton steps (Goldschmidt's for the
// root): within an ulp or two of the correctly rounded values, in about half
// the instructions of the IEEE sequences, which scale, classify and fix up
// every operand.  For the FP32 march's f64 finish: its decisions need f64
// resolution, not the reference's rounding of every operation (that is the
// FP64 mode's, which keeps the exact forms).
__device__ __forceinline__ double sqrt_fast64(double x) {
    if (!(x >= 0x1p-900 && x <= 0x1p900)) return sqrt(x);  // 0, tiny, huge, inf, NaN: the exact form
    const double y = __builtin_amdgcn_rsq(x);
    double g = x * y, h = 0.5 * y;
    const double r = __builtin_fma(-g, h, 0.5);
    g = __builtin_fma(g, r, g);
    h = __builtin_fma(h, r, h);
    const double d = __builtin_fma(-g, g, x);
    return __builtin_fma(d, h, g);
}

// 1 / d for a normal positive d below 2^900
__device__ __forceinline__ double rcp_fast64(double d) {
    double y = __builtin_amdgcn_rcp(d);
    double e = __builtin_fma(-d, y, 1.0);
    y = __builtin_fma(y, e, y);
    e = __builtin_fma(-d, y, 1.0);
    return __builtin_fma(y, e, y);
}

// int_pow8_64 in the form of p8f (the Chebyshev identities, the fourth
// powers from r2 = |v|^2), with FMAs and the refined hardware root and
// reciprocal: the finish's form (the divisor R^4 + 1e-40 lies in
// [1e-40, 2^16] for an orbit that has not escaped)
__device__ __forceinline__ void int_pow8_64_fast(double zmul, double v[4], const double j[4], double r2) {
    const double x = v[0], y = v[1], z = v[2];
    const double xx = x * x, yy = y * y, zz = z * z;
    const double R = __builtin_fma(x, x, yy), rr = R * R, zr = zz * R;
    const double t = __builtin_fma(-6.0, zr, __builtin_fma(zz, zz, rr)) * (zz - R) * sqrt_fast64(R);
    v[2] = __builtin_fma(-(t * z), 8.0 * zmul, j[2]);
    const double s2 = r2 * r2;
    const double a = __builtin_fma(s2, s2, zr * __builtin_fma(128.0, zr, -32.0 * s2)) *
                     rcp_fast64(__builtin_fma(rr, rr, 1e-40));
    const double p2 = xx * yy;
    v[1] = __builtin_fma((xx - yy) * __builtin_fma(-8.0, p2, rr) * (x * y), 8.0 * a, j[1]);
    v[0] = __builtin_fma(a, __builtin_fma(rr, rr, p2 * __builtin_fma(128.0, p2, -32.0 * rr)), j[0]);
}

// amazing_box (src/formulas.rs:713-736)
__device__ __forceinline__ void box64(double scale, double min_r, double fold, double v[4], const double j[4],
                                      bool ade) {
    const double x = bfold64(v[0], fold), y = bfold64(v[1], fold), z = bfold64(v[2], fold);
    const double r = z * z + y * y + x * x;
    const double m = fmax(min_r, 1e-40);
    const double sqr_min_r = m * m;
    double mul;
    if (r < sqr_min_r)
        mul = scale / sqr_min_r;
    else if (1.0 < r)
        mul = scale;
    else
        mul = scale / r;
    if (ade) v[3] *= mul;
    v[2] = z * mul + j[2];
    v[1] = y * mul + j[1];
    v[0] = x * mul + j[0];
}

// box64 with FMAs and the refined hardware reciprocal (the finish's form;
// r lies in [min_r^2, 1] where it divides)
__device__ __forceinline__ void box64_fast(double scale, double min_r, double fold, double v[4], const double j[4],
                                           bool ade) {
    const double x = bfold64(v[0], fold), y = bfold64(v[1], fold), z = bfold64(v[2], fold);
    const double r = __builtin_fma(x, x, __builtin_fma(y, y, z * z));
    const double m = fmax(min_r, 1e-40);
    const double sqr_min_r = m * m;
    double mul;
    if (r < sqr_min_r)
        mul = scale * rcp_fast64(sqr_min_r);
    else if (1.0 < r)
        mul = scale;
    else
        mul = scale * rcp_fast64(r);
    if (ade) v[3] *= mul;
    v[2] = __builtin_fma(z, mul, j[2]);
    v[1] = __builtin_fma(y, mul, j[1]);
    v[0] = __builtin_fma(x, mul, j[0]);
}

// int_pow2_scale (src/formulas.rs:739-752), the bulb of Bulbox
__device__ __forceinline__ void pow2_scale64(double scaling, double v[4], const double j[4]) {
    const double s = 1.0 / scaling;
    const double x = v[0] * s, y = v[1] * s, z = v[2] * s;
    const double xx = x * x, yy = y * y, r = xx + yy;
    v[2] = -((sqrt(r) * z) * 2.0) * scaling + j[2];
    const double a = (r - z * z) / r;
    v[0] = (xx - yy) * a * scaling + j[0];
    v[1] = (x * y * a) * 2.0 * scaling + j[1];
}

// Bulbox, HybridSuperCube2 (src/formulas.rs:596-627)
// box scale, Min R, fold; bulb scaling, R threshold, R threshold 2
__device__ __forceinline__ void bulbox64(double scale, double min_r, double fold, double bulb, double rth,
                                         double rth2, double v[4], const double j[4], double rout) {
    const double rt1 = rth * rth, rt2 = rth2 * rth2;
    if (!(rout < rt1)) {
        box64(scale, min_r, fold, v, j, false);
        return;
    }
    if (rout < rt2) {
        pow2_scale64(bulb, v, j);
        return;
    }
    const double m = (rout - rt2) / (rt1 - rt2);
    const double in0 = v[0], in1 = v[1], in2 = v[2];
    pow2_scale64(bulb, v, j);
    const double b0 = v[0], b1 = v[1], b2 = v[2];
    v[0] = in0;
    v[1] = in1;
    v[2] = in2;
    box64(scale, min_r, fold, v, j, false);
    double r1 = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    const double m2 = 1.0 - m;
    const double r2 = r1 * m + m2 * sqrt(b0 * b0 + b1 * b1 + b2 * b2);
    v[0] = v[0] * m + m2 * b0;
    v[1] = v[1] * m + m2 * b1;
    v[2] = v[2] * m + m2 * b2;
    r1 = r2 / sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2] + 1e-40);
    v[0] *= r1;
    v[1] *= r1;
    v[2] *= r1;
}

// real_power, HybridFloatPow (src/formulas.rs:755-766): angles by atan2,
// the radius from the stored squared radius
__device__ __forceinline__ void real_power64(double pw, double zmul, double v[4], const double j[4], double rout) {
    const double x = v[0], y = v[1], z = v[2];
    double st, ct, sp, cp;
    sincos(atan2(y, x) * pw, &st, &ct);
    sincos(atan2(z, sqrt(x * x + y * y)) * pw, &sp, &cp);
    const double r = pow(rout, pw * 0.5);
    v[0] = ct * cp * r + j[0];
    v[1] = st * cp * r + j[1];
    v[2] = sp * r * zmul + j[2];
}

// aexion_c (src/formulas.rs:644-703): a real power bulb step with the
// angles from the y axis; with "rotate C" the constant is turned by a power
// of its own angles (or those of Z - C, mode 16)
// power, Z multiplier, power C, Cz multiplier, the switches and mode
__device__ __forceinline__ void aexion64(double pw, double z_mul, double power_c, double cz_mul, int flags,
                                         double v[4], double j[4]) {
    const double x = v[0], y = v[1], z = v[2];
    double st, ct, sp, cp;
    sincos(atan2(sqrt(x * x + z * z), y) * pw, &st, &ct);
    sincos(atan2(z, x) * pw, &sp, &cp);
    const double r1 = exp(log(x * x + y * y + z * z) * pw * 0.5);
    v[0] = cp * ct * r1 + j[0];
    v[2] = st * r1 * z_mul + j[2];
    v[1] = ct * sp * r1 + j[1];
    if (!(flags & 1)) return;
    const int mode = flags >> 8;
    double pd = power_c;
    if (flags & 4) {
        const double d0 = v[0] - j[0], d1 = v[1] - j[1], d2 = v[2] - j[2];
        pd *= sqrt(d0 * d0 + d1 * d1 + d2 * d2);
    }
    const double r = sqrt(j[0] * j[0] + j[1] * j[1] + j[2] * j[2]);
    double a = (mode & 16) ? v[1] - j[1] : j[1];
    double b = (mode & 16) ? v[2] - j[2] : j[2];
    const double c = (mode & 16) ? v[0] - j[0] : j[0];
    int tested = 0, other = 1;  // the component "conditional phi" tests: 0 x, 1 y, 2 z
    if (mode & 8) {
        const double t = a;
        a = b;
        b = t;
        other = 2;
    }
    const double s = sqrt(b * b + c * c);
    double th = (mode & 1) ? atan2(a, s) : atan2(s, a);
    double ph;
    if (mode & 2) {
        tested = 3 - other;
        ph = atan2(c, b);
    } else {
        ph = atan2(b, c);
    }
    if (mode & 4) {
        const double t = th;
        th = ph;
        ph = t;
        tested = other;
    }
    if (flags & 2) {
        const double t = tested == 0 ? v[0] : (tested == 1 ? v[1] : v[2]);
        if (!signbit(t)) ph = -ph;
    }
    double sa, ca, sb, cb;
    sincos(ph * pd, &sa, &ca);
    sincos(th * pd, &sb, &cb);
    j[0] = ca * cb * r;
    j[2] = sb * r * cz_mul;
    j[1] = cb * sa * r;
}

// Formula::iterate via call_slot (src/iteration.rs:325-345,
// src/formulas.rs:576-641).  The power-8 and box kinds keep their two
// formulas inline; K_GEN selects among all built-ins.
// FAST: the power-8 and box kinds' two formulas in their fast forms (the
// finish of the FP32 march); K_GEN keeps the exact forms.
template <int KIND, bool FAST = false>
__device__ __forceinline__ void call_slot64(const DevP &P, int n, It64 &s) {
    double *v = s.v;
    // the options as values (a pointer into the kernel's parameter block
    // would make the compiler copy the block to scratch memory)
    const double o[3] = {P.opt[n][0], P.opt[n][1], P.opt[n][2]};
    const int kind = P.slot_kind[n];
    if (KIND != K_GEN) {
        // (K_P8 and K_BOX hold one formula in every counted slot: the other's
        // code is left out)
        if (KIND == K_P8 || (KIND == K_HYB && kind == MB3D_HIP_FORMULA_INTPOW)) {
            if (P.is_custom_de) {  // the analytic power-8 extension: dr <- 8 r^7 dr + 1
                // (s.rout is v's squared radius, v[0]^2 + v[1]^2 + v[2]^2 in
                // this order: start64 and run64 keep it after every counted
                // iteration, and these kinds count every slot)
                const double r2 = s.rout;
                if (FAST)
                    s.deriv1 = __builtin_fma(8.0 * (r2 * r2 * r2 * sqrt_fast64(r2)), s.deriv1, 1.0);
                else
                    s.deriv1 = 8.0 * (r2 * r2 * r2 * sqrt(r2)) * s.deriv1 + 1.0;
            }
            if (FAST)
                int_pow8_64_fast(o[0], v, s.j, s.rout);
            else
                int_pow8_64(o[0], v, s.j);
            s.p8++;
        } else {
            if (FAST)
                box64_fast(o[0], o[1], o[2], v, s.j, P.is_custom_de != 0);
            else
                box64(o[0], o[1], o[2], v, s.j, P.is_custom_de != 0);
            s.box++;
        }
        return;
    }
    switch (kind) {
        case MB3D_HIP_FORMULA_INTPOW:
            int_pow64(P.slot_iarg[n], o[0], v, s.j);
            s.p8++;
            return;
        case MB3D_HIP_FORMULA_AMAZINGBOX:
            box64(o[0], o[1], o[2], v, s.j, P.slot_ade[n] != 0);
            s.box++;
            return;
        case MB3D_HIP_FORMULA_REALPOWER: real_power64(o[0], o[1], v, s.j, s.rout); break;
        case MB3D_HIP_FORMULA_QUATERNION: {  // src/formulas.rs:580-586
            const double x = v[0], y = v[1], z = v[2], w = v[3];
            v[0] = x * x - y * y - z * z - w * w + s.j[0];
            v[1] = 2.0 * (y * x + z * w) + s.j[1];
            v[2] = 2.0 * (z * x + o[0] * y * w) + s.j[2];
            v[3] = 2.0 * (w * x + y * z) + o[1] + s.j[3];
            break;
        }
        case MB3D_HIP_FORMULA_TRICORN: {  // src/formulas.rs:587-592
            const double x = v[0], y = v[1], z = v[2];
            v[2] = z * x * o[0] + s.j[2] * o[1];
            v[0] = x * x - (y * y + z * z) + s.j[0];
            v[1] = 2.0 * (y * x) + s.j[1];
            break;
        }
        case MB3D_HIP_FORMULA_BULBOX:
            bulbox64(o[0], o[1], o[2], P.opt2[n][0], P.opt2[n][1], P.opt2[n][2], v, s.j, s.rout);
            break;
        case MB3D_HIP_FORMULA_FOLDINGINTPOW:  // src/formulas.rs:629-634
            v[0] = bfold64(v[0], o[1]);
            v[1] = bfold64(v[1], o[1]);
            v[2] = bfold64(v[2], o[1]);
            int_pow64(P.slot_iarg[n], o[0], v, s.j);
            break;
        default: aexion64(o[0], o[1], o[2], P.opt2[n][0], P.slot_iarg[n], v, s.j); break;
    }
    s.oth++;
}

// calc_smooth_iterations (src/iteration.rs:140-154)
__device__ __forceinline__ float smooth64(const DevP &P, const It64 &s, int n) {
    if (s.rout <= 1.0) return (float)s.it_result;
    if (s.rold < 1.0) {
        const double d = log(log(s.rout) * 0.5) * (double)P.fhln[n];
        return (float)((double)s.it_result + (double)P.ln_rstop - d);
    }
    const double d = log(0.5 * log(s.rout));
    const double e = d - log(0.5 * log(s.rold));
    return (float)(((double)P.ln_rstop - d) / (e + 1e-100) + (double)s.it_result);
}

// run_ipol (src/iteration.rs:347-401): the interpolation hybrid of slots 0
// and 1 -- both applied to the same vector, the results and their lengths
// blended by the weights; with an analytic DE the derivative in w (3D)
// blends too.  Returns the slot the smooth iterations use (0).
__device__ __forceinline__ int run_ipol64(const DevP &P, It64 &s, int max_it, double rstop) {
    const double s1 = (double)P.ipol_w[0], s2 = (double)P.ipol_w[1];
    const bool four = P.four_d != 0, de = P.is_custom_de != 0;
    s.it_result = 0;
    for (;;) {
        s.rold = s.rout;
        const double y0 = s.v[0], y1 = s.v[1], y2 = s.v[2], y3 = s.v[3];
        call_slot64<K_GEN>(P, 0, s);
        const double a0 = s.v[0], a1 = s.v[1], a2 = s.v[2], a3 = s.v[3];
        s.v[0] = y0;
        s.v[1] = y1;
        s.v[2] = y2;
        s.v[3] = y3;
        call_slot64<K_GEN>(P, 1, s);
        const double x = s.v[0], y = s.v[1], z = s.v[2], w = s.v[3];
        double la, lb;
        if (four) {
            la = sqrt(a0 * a0 + a1 * a1 + a2 * a2 + a3 * a3);
            lb = sqrt(x * x + y * y + z * z + w * w);
        } else {
            la = sqrt(a0 * a0 + a1 * a1 + a2 * a2);
            lb = sqrt(x * x + y * y + z * z);
        }
        const double len = la * s1 + lb * s2;
        double v0 = a0 * s1 + x * s2, v1 = a1 * s1 + y * s2, v2 = a2 * s1 + z * s2, v3 = a3 * s1 + w * s2;
        if (de && !four) v3 = fabs(a3) * s1 + fabs(w) * s2;
        const double l2 = four ? v0 * v0 + v1 * v1 + v2 * v2 + v3 * v3 : v0 * v0 + v1 * v1 + v2 * v2;
        const double f = len / sqrt(l2 + 1e-40);
        s.v[0] = v0 * f;
        s.v[1] = v1 * f;
        s.v[2] = v2 * f;
        s.v[3] = four ? v3 * f : v3;
        s.it_result += 1;
        s.rout = len * len;
        if (s.rout < s.otrap) s.otrap = s.rout;
        if (s.it_result >= max_it || s.rout > rstop) return 0;
    }
}

// One hybrid part (HybridPart, src/calc.rs:137-150; set_part,
// src/calc.rs:1470-1490): its slots, its DE option and scale, 3D or 4D.
// A picture without a DE combination has one part.
struct Part64 {
    int start, end, repeat, de_option, custom, four;
    double dds;
};

template <int KIND>
__device__ __forceinline__ Part64 part1_of(const DevP &P) {
    return {0, P.end_to, P.repeat_from, P.de_option, P.is_custom_de, KIND == K_GEN && P.four_d, (double)P.d_de_scale};
}

// run_with (src/iteration.rs:268-315), counted slots, 3D or 4D squared
// radius; returns the index of the last slot used
template <int KIND, bool FAST = false>
__device__ __forceinline__ int run64(const DevP &P, const Part64 &pt, It64 &s, int max_it, double rstop) {
    if (KIND == K_GEN && P.ipol) return run_ipol64(P, s, max_it, rstop);
    int n = pt.start;
    int btmp = P.slot_iters[n];
    s.it_result = 0;
    uint32_t uncounted_guard = 0;
    for (;;) {
        s.rold = s.rout;
        int guard = 0;
        while (btmp <= 0) {
            n += 1;
            if (n > pt.end) n = pt.repeat;
            btmp = P.slot_iters[n];
            if (++guard > 64) return n;
        }
        call_slot64<KIND, FAST>(P, n, s);
        btmp -= 1;
        if (KIND == K_GEN && P.slot_unc[n]) {  // a slot that does not count (src/iteration.rs:294-300)
            if (++uncounted_guard > 100000) return n;
            continue;
        }
        s.it_result += 1;
        if (KIND == K_GEN && pt.four)
            s.rout = s.v[0] * s.v[0] + s.v[1] * s.v[1] + s.v[2] * s.v[2] + s.v[3] * s.v[3];
        else if (FAST && KIND != K_GEN)
            s.rout = __builtin_fma(s.v[2], s.v[2], __builtin_fma(s.v[1], s.v[1], s.v[0] * s.v[0]));
        else
            s.rout = s.v[0] * s.v[0] + s.v[1] * s.v[1] + s.v[2] * s.v[2];
        if (s.rout < s.otrap) s.otrap = s.rout;
        if (s.it_result >= max_it || s.rout > rstop) return n;
    }
}

// The start of an evaluation at c: hybrid_3d / hybrid_3d_de
// (src/iteration.rs:521-552), the Julia seed as the constant in Julia
// mode, J4 kept in j[3]; the 4D hybrids rotate c into 4D first (start_4d,
// rotate_4dex: src/iteration.rs:417-428, src/math.rs:209-215).
template <int KIND>
__device__ __forceinline__ void start64(const DevP &P, bool four, It64 &s, double cx, double cy, double cz) {
    if (KIND == K_GEN && four) {
        for (int i = 0; i < 4; i++) {
            s.v[i] = cx * (double)P.sm4[3 * i] + cy * (double)P.sm4[3 * i + 1] + cz * (double)P.sm4[3 * i + 2];
            s.j[i] = P.do_julia ? P.ju[i] : s.v[i];
        }
        s.rout = s.v[0] * s.v[0] + s.v[1] * s.v[1] + s.v[2] * s.v[2] + s.v[3] * s.v[3];
        s.otrap = s.rout;
        return;
    }
    const bool julia = KIND == K_GEN && P.do_julia;
    s.j[0] = julia ? P.ju[0] : cx;
    s.j[1] = julia ? P.ju[1] : cy;
    s.j[2] = julia ? P.ju[2] : cz;
    s.j[3] = KIND == K_GEN ? P.ju[3] : 0.0;
    s.v[0] = cx;
    s.v[1] = cy;
    s.v[2] = cz;
    s.v[3] = 0.0;
    s.rout = cx * cx + cy * cy + cz * cz;
    s.otrap = s.rout;
}

// The result of one f64 DE evaluation.
struct De64 {
    double de, rout, otrap;
    double v[3], rold;  // the orbit's end and the radius before it (do_color)
    int it_result;
    int max_its;       // max_its_result: the ruling part's iteration limit
    int max_it_after;  // the iteration limit the evaluation leaves (set_part)
    float smooth_it;
    uint32_t p8, box, oth;
};

// calc_de_part without dIFS (src/calc.rs:1495-1559) for one part:
// hybrid_3d_de (src/iteration.rs:539-585) for the analytic options --
// option 32 being the textbook 0.5 r ln r / dr of the power-8 extension --
// or the alternating hybrid (3D or 4D) four times for the numeric gradient.
// Arguments and result in registers.
// DEM: the DE mode when the kernel knows it (DEM_P8 and DEM_BOX are
// analytic: the numeric gradient's code is left out), DEM_NUMERIC when the
// part's own flag decides.
template <int KIND, int DEM = DEM_NUMERIC, bool FAST = false>
__device__ __forceinline__ De64 de64_part(const DevP &P, const Part64 &pt, double cx, double cy, double cz, int max_it,
                                       int calc_sit, float ms, double rstop = -1.0) {
    // (rstop: the bailout of the evaluation, P.d_rstop unless given;
    // do_color_on_it asks for rstop3d)
    const double rs = rstop >= 0.0 ? rstop : P.d_rstop;
    De64 o;
    It64 s;
    s.p8 = s.box = s.oth = 0;
    s.deriv1 = 1.0;
    s.rold = 0.0;
    double result;
    o.smooth_it = 0.0f;
    start64<KIND>(P, pt.four, s, cx, cy, cz);
    if (DEM != DEM_NUMERIC || pt.custom) {
        if ((pt.de_option & 0x38) == 32)
            s.deriv1 = 1.0;
        else if ((pt.de_option & 0x38) == 16)
            s.v[3] = s.rout;
        else
            s.v[3] = 1.0;
        const int n = run64<KIND, FAST>(P, pt, s, max_it, rs);
        if (FAST) {
            // (the refined reciprocal where its operand allows it, the
            // IEEE quotient otherwise: the derivative can reach 2^900)
            const double den = (pt.de_option & 0x38) == 32 ? s.deriv1 : fabs(s.v[3]);
            const double inv = den >= 0x1p-900 && den <= 0x1p900 ? rcp_fast64(den) : 1.0 / den;
            result = (pt.de_option & 0x38) == 32 ? sqrt_fast64(s.rout) * 0.25 * log(s.rout) * inv
                                                 : sqrt_fast64(s.rout) * inv;
        } else if ((pt.de_option & 0x38) == 32) {
            result = sqrt(s.rout) * 0.25 * log(s.rout) / s.deriv1;
        } else {
            result = sqrt(s.rout) / fabs(s.v[3]);
        }
        if (calc_sit) o.smooth_it = smooth64(P, s, n);
        result = result * pt.dds;
        o.it_result = s.it_result;
        o.rout = s.rout;
        o.otrap = s.otrap;
    } else {
        const int n = run64<KIND, FAST>(P, pt, s, max_it, rs);
        if (calc_sit) o.smooth_it = smooth64(P, s, n);
        o.it_result = s.it_result;
        o.rout = s.rout;
        o.otrap = s.otrap;
        if (s.rout < 1e-200) {
            result = 0.0;
        } else {
            const double buf_rout = s.rout;
            const int it0 = s.it_result;
            const double off = (double)P.de_offset;
            double g[3];
            for (int k = 0; k < 3; k++) {
                start64<KIND>(P, pt.four, s, k == 0 ? cx + off : cx, k == 1 ? cy + off : cy, k == 2 ? cz + off : cz);
                run64<KIND, FAST>(P, pt, s, it0, P.rstop3d);
                g[k] = (buf_rout - s.rout) * (buf_rout - s.rout);
            }
            result = buf_rout * log(buf_rout) * pt.dds / (sqrt(g[0] + g[1] + g[2]) + (double)P.de_offset006);
            o.otrap = s.otrap;  // the last pass's, as the reference leaves it
        }
    }
    const double mn = (double)ms * 0.25;
    if (result < mn || isnan(result)) result = mn;
    o.de = result;
    o.v[0] = s.v[0];
    o.v[1] = s.v[1];
    o.v[2] = s.v[2];
    o.rold = s.rold;
    o.max_its = max_it;
    o.max_it_after = max_it;
    o.p8 = s.p8;
    o.box = s.box;
    o.oth = s.oth;
    return o;
}

// calc_de_fractal with a DE combination (src/calc.rs:1352-1456): the first
// part at c, the second at its own place (and inverted for kind 3), then the
// minimum, maximum or a smooth minimum of the two; the iteration count,
// trap and smooth iterations of the ruling part (blended for the smooth
// minima).  Afterwards the iteration limit is the first part's.
__device__ __forceinline__ De64 de64_comb(const DevP &P, double cx, double cy, double cz, int max_it, int calc_sit,
                                       float ms) {
    const De64 a = de64_part<K_GEN>(P, part1_of<K_GEN>(P), cx, cy, cz, max_it, calc_sit, ms);
    double qx = cx, qy = cy, qz = cz;
    if (P.placed) {  // Place::apply: m (p - offset)
        const double d0 = cx - P.place_off[0], d1 = cy - P.place_off[1], d2 = cz - P.place_off[2];
        qx = P.place_m[0][0] * d0 + P.place_m[0][1] * d1 + P.place_m[0][2] * d2;
        qy = P.place_m[1][0] * d0 + P.place_m[1][1] * d1 + P.place_m[1][2] * d2;
        qz = P.place_m[2][0] * d0 + P.place_m[2][1] * d1 + P.place_m[2][2] * d2;
    }
    const Part64 p2 = {P.start2, P.end2, P.repeat2, P.de_option2, P.custom2, P.four_d2, (double)P.d_de_scale2};
    De64 o = de64_part<K_GEN>(P, p2, qx, qy, qz, P.max_it2, calc_sit, ms);
    double rst = o.de;
    const double msd = (double)ms;
    if (P.decomb == 3) {  // the second part inverted (calc_de_part's `invert`)
        rst = msd * 4.0 - rst * 3.0;
        if (rst >= msd) o.it_result -= 1;
    }
    if (P.placed) rst *= P.place_scale;
    double result = a.de;
    bool first_rules = false;
    if (P.decomb == 1) {
        if (rst >= result)
            first_rules = true;
        else
            result = rst;
    } else if (P.decomb <= 3) {
        if (rst > result)
            result = rst;
        else
            first_rules = true;
    } else {
        if (!(rst < result)) {
            o.it_result = a.it_result;
            o.max_its = a.max_its;
        }
        const double sm = P.smooth;
        if (P.decomb == 4) {
            result = fmin(rst - fmax(sm - result, 0.0), result - fmax(sm - rst, 0.0));
        } else {
            const double rd1 = fmax(sm - result - msd, 0.0), d2 = fmax(sm - rst - msd, 0.0);
            result = fmin(rst - rd1 * (sm - d2) / sm, result - d2 * (sm - rd1) / sm);
        }
        const double r2 = fabs(rst), w = fabs(result);
        const double f = 1.0 / (w + r2 + 1e-10);
        o.smooth_it = (float)(((double)o.smooth_it * w + (double)a.smooth_it * r2) * f);
        o.otrap = (o.otrap * w + a.otrap * r2) * f;
    }
    if (first_rules) {
        o.it_result = a.it_result;
        o.smooth_it = a.smooth_it;
        o.otrap = a.otrap;
        o.max_its = a.max_its;
    }
    o.de = result;
    o.max_it_after = P.max_it;
    o.p8 += a.p8;
    o.box += a.box;
    o.oth += a.oth;
    return o;
}

// FAST: the power-8 and box kinds' iterates in their fast forms
// (int_pow8_64_fast, box64_fast) for the FP32 march's finish and cone rays
template <int KIND, int DEM = DEM_NUMERIC, bool FAST = false>
__device__ __forceinline__ De64 de64(const DevP &P, double cx, double cy, double cz, int max_it, int calc_sit,
                                  float ms, double rstop = -1.0) {
    if (KIND == K_GEN && P.decomb) return de64_comb(P, cx, cy, cz, max_it, calc_sit, ms);
    return de64_part<KIND, KIND == K_GEN ? DEM_NUMERIC : DEM, FAST && KIND != K_GEN>(P, part1_of<KIND>(P), cx, cy, cz,
                                                                                     max_it, calc_sit, ms, rstop);
}

// mand_function (src/calc.rs:1277-1285, Iteration::hybrid_3d,
// src/iteration.rs:521-535): the plain iteration at c, no DE, with the
// smooth iteration count; the result in a De64 (its `de` unset)
template <int KIND>
__device__ __forceinline__ De64 mand64(const DevP &P, double cx, double cy, double cz, int max_it) {
    De64 o;
    It64 s;
    s.p8 = s.box = s.oth = 0;
    s.deriv1 = 1.0;
    s.rold = 0.0;
    const Part64 pt = part1_of<KIND>(P);
    start64<KIND>(P, pt.four, s, cx, cy, cz);
    const int n = run64<KIND>(P, pt, s, max_it, P.d_rstop);
    o.smooth_it = smooth64(P, s, n);
    o.de = 0.0;
    o.rout = s.rout;
    o.otrap = s.otrap;
    o.v[0] = s.v[0];
    o.v[1] = s.v[1];
    o.v[2] = s.v[2];
    o.rold = s.rold;
    o.it_result = s.it_result;
    o.max_its = o.max_it_after = max_it;
    o.p8 = s.p8;
    o.box = s.box;
    o.oth = s.oth;
    return o;
}

// The reference's Marcher (src/calc.rs:1143-1170) for one ray in f64: the
// position, direction and depth, the DE stop, and the values of its last DE
// evaluation that the decisions read.
struct Ray64 {
    double c[3], v[3];
    double mzz;
    float ms;  // ms_de_stop
    int max_it;
    bool calc_sit;
    // the last DE
    double de, rout, otrap;
    int it, max_its;  // it_result, max_its_result
    float smooth_it;

    __device__ __forceinline__ void reset(const DevP &P) {
        mzz = 0.0;
        ms = P.de_stop;
        max_it = P.max_it;
        calc_sit = false;
        de = rout = otrap = 0.0;
        it = 0;
        max_its = P.max_it;
        smooth_it = 0.0f;
    }
    // mzz += d; step(d)
    __device__ __forceinline__ void move(double d) {
        mzz += d;
        c[0] += v[0] * d;
        c[1] += v[1] * d;
        c[2] += v[2] * d;
    }
    // update_de_stop (src/calc.rs:1698-1703)
    __device__ __forceinline__ void upd(const DevP &P) {
        ms = (float)((double)P.de_stop * (1.0 + mzz * (double)P.de_stop_factor));
    }
    // takes the result of de64 at c
    __device__ __forceinline__ void take(const De64 &r, Cnt &cnt) {
        de = r.de;
        it = r.it_result;
        rout = r.rout;
        otrap = r.otrap;
        if (calc_sit) smooth_it = r.smooth_it;
        max_its = r.max_its;
        max_it = r.max_it_after;
        cnt.de64++;
        cnt.p8_64 += r.p8;
        cnt.box_64 += r.box;
        cnt.oth_64 += r.oth;
    }
    // calc_zpos_and_rough, smooth_normals = 0 (src/calc.rs:2194-2205)
    __device__ __forceinline__ uint32_t zpos_fine(const DevP &P) const {
        const double zz = fmax(mzz, 0.0);
        const double vr = rte(P.zc_mul * (sqrt(zz * P.zcorr + 1.0) - 1.0));
        const int64_t vv = (int64_t)vr;
        int64_t itmp = 8388352 - vv;
        itmp = itmp < 0 ? 0 : (itmp > 8388352 ? 8388352 : itmp);
        return (uint32_t)itmp << 8;
    }
};

// The surface refinements as resumable loops: each call runs until the loop
// needs its next DE at the ray's position (returns true) or ends (false).
//   RK_BS   bin_search     (src/calc.rs:1706-1735)
//   RK_BSI  bin_search_it  (src/calc.rs:1737-1792)
//   RK_SP   the bisection of surface_point (src/calc.rs:2173-2185)
enum RefKind : int { RK_BS = 0, RK_BSI = 1, RK_SP = 2 };

struct Refine {
    int kind, itmp, saved_max, sub;
    double dt1, d, dmul, thr;
    float yp, last_si, last_dif;
    bool first;

    __device__ __forceinline__ bool bs_head(Ray64 &R, const DevP &P) {
        if (fabs(d - (double)R.ms) > 0.001) {
            R.move(dt1);
            R.upd(P);
            itmp -= 1;
            return itmp > 0;
        }
        return false;
    }
    __device__ __forceinline__ bool start_bs(Ray64 &R, const DevP &P, double de, double last_step_width, int steps) {
        kind = RK_BS;
        itmp = steps;
        dt1 = last_step_width * -0.5;
        d = de;
        return bs_head(R, P);
    }
    __device__ __forceinline__ bool start_bsi(Ray64 &R, int steps) {
        kind = RK_BSI;
        yp = (float)R.max_it - 0.99f;
        saved_max = R.max_it;
        R.max_it += 1;
        itmp = steps;
        R.calc_sit = true;
        dt1 = 0.0;
        dmul = 1.0;
        first = true;
        last_si = last_dif = 0.0f;
        sub = 0;
        R.move(dt1);
        return true;
    }
    __device__ __forceinline__ bool sp_head(Ray64 &R, const DevP &P) {
        if (itmp > 0 && fabs(d - (double)R.ms) > thr) {
            R.move(dt1);
            R.upd(P);
            itmp -= 1;
            return itmp > 0;
        }
        return false;
    }
    __device__ __forceinline__ bool start_sp(Ray64 &R, const DevP &P, double de, double dt, int steps, double th) {
        kind = RK_SP;
        itmp = steps;
        d = de;
        dt1 = dt;
        thr = th;
        return sp_head(R, P);
    }
    // after the DE the loop asked for
    __device__ __forceinline__ bool after(Ray64 &R, const DevP &P) {
        if (kind == RK_BS) {
            d = R.de;
            if (R.it >= R.max_its) {
                dt1 = -fabs(dt1);
            } else {
                const double f = d - (double)R.ms;
                dt1 = f < 0.0 ? fabs(dt1) * -0.55 : fabs(dt1) * 0.55;
            }
            return bs_head(R, P);
        }
        if (kind == RK_SP) {
            d = R.de;
            dt1 = d < (double)R.ms ? fabs(dt1) * -0.55 : fabs(dt1) * 0.55;
            return sp_head(R, P);
        }
        // bin_search_it: the main probe, then the probe 0.001 nearer
        if (sub == 0) {
            if (!first && last_dif < fabsf(yp - R.smooth_it)) {
                R.move(-dt1);
                R.smooth_it = last_si;
                dmul *= dt1 > 0.0 ? 0.5 : 0.7;
            }
            last_dif = fabsf(yp - R.smooth_it);
            last_si = R.smooth_it;
            if (!(R.smooth_it > (float)R.max_it - 0.1f)) {
                R.move(-0.001);
                sub = 1;
                return true;
            }
            dt1 = -3.0;
        } else {
            const float r = last_si - R.smooth_it;
            if (fabsf(r) < 1e-30f)
                dt1 = (last_si < R.smooth_it ? 1.0 : 0.0) - 0.5;
            else if (r < 0.0f)
                dt1 = (double)((yp - R.smooth_it) / (r * 500.0f));
            else
                dt1 = (double)((yp - R.smooth_it) / (r * 1000.0f));
            if (dt1 > 4.0)
                dt1 = sqrt(dt1) * 2.0;
            else if (dt1 < -9.0)
                dt1 = sqrt(-dt1) * -3.0;
            dt1 = dt1 * dmul + 0.0005;
            sub = 0;
        }
        first = false;
        itmp -= 1;
        if (itmp < 0) {
            R.max_it = saved_max;
            return false;
        }
        R.move(dt1);
        return true;
    }
};

// The ray of pixel (x, y) (calc_vgrads_fov + calc_start_pos,
// src/calc.rs:1561-1575,1673-1690; build_view_vector_dfov, rotate_vector_reverse
// src/math.rs:63-69,127-133), Common optic.
__device__ __forceinline__ void pixel_ray(const DevP &P, int x, int y, double c[3], double v[3]) {
    const double cafy = ((double)y / (double)P.height - 0.5) * P.fov_y;
    const double cafx = (double)((P.fovx_off - (float)(x + 1)) * P.fovx_mul);
    double sx, cx, sy, cy;
    sincos(cafy, &sx, &cx);
    sincos(cafx, &sy, &cy);
    double w0 = -sy, w1 = sx, w2 = cx * cy;
    const double nrm = 1.0 / sqrt(w0 * w0 + w1 * w1 + w2 * w2);
    w0 *= nrm;
    w1 *= nrm;
    w2 *= nrm;
    for (int k = 0; k < 3; k++) v[k] = w0 * P.vg[0][k] + w1 * P.vg[1][k] + w2 * P.vg[2][k];
    const double fx = x, fy = y;
    for (int k = 0; k < 3; k++) c[k] = P.ystart[k] + P.vg[0][k] * fx + P.vg[1][k] * fy;
}

// The ray tables of a frame: sincos of calc_vgrads_fov's two angles, which
// depend on the column (cafx) and on the row (cafy) alone.  Entry rt_base +
// x holds (sin, cos) of column x's angle, entry rt_base + width + yr those
// of row row0 + yr.  Filled by k_ray_tab with the same sincos calls as
// pixel_ray, so that pixel_ray_tab's rays are pixel_ray's bit for bit.
__device__ __forceinline__ void ray_tab_entry(const DevP &P, int i, double2 &out) {
    double a;
    if (i < P.width) {
        a = (double)((P.fovx_off - (float)(i + 1)) * P.fovx_mul);
    } else {
        const int y = P.row0 + (i - P.width);
        a = ((double)y / (double)P.height - 0.5) * P.fov_y;
    }
    sincos(a, &out.x, &out.y);
}

// pixel_ray with the frame's ray tables (`tab`): the view vector from the
// column's and the row's sincos, normalised and rotated as in pixel_ray
__device__ __forceinline__ void pixel_ray_tab(const DevP &P, const double2 *__restrict__ tab, int x, int yr, double c[3],
                                              double v[3]) {
    const double2 col = tab[P.rt_base + x];
    const double2 row = tab[P.rt_base + P.width + yr];
    const double sx = row.x, cx = row.y, sy = col.x, cy = col.y;
    double w0 = -sy, w1 = sx, w2 = cx * cy;
    const double nrm = 1.0 / sqrt(w0 * w0 + w1 * w1 + w2 * w2);
    w0 *= nrm;
    w1 *= nrm;
    w2 *= nrm;
    for (int k = 0; k < 3; k++) v[k] = w0 * P.vg[0][k] + w1 * P.vg[1][k] + w2 * P.vg[2][k];
    const double fx = x, fy = P.row0 + yr;
    for (int k = 0; k < 3; k++) c[k] = P.ystart[k] + P.vg[0][k] * fx + P.vg[1][k] * fy;
}

// pixel_ray_tab's start point alone (the same operations: the same bits)
__device__ __forceinline__ void pixel_start(const DevP &P, int x, int yr, double c[3]) {
    const double fx = x, fy = P.row0 + yr;
    for (int k = 0; k < 3; k++) c[k] = P.ystart[k] + P.vg[0][k] * fx + P.vg[1][k] * fy;
}

// The ray of the picture's point (fx, fy), fractions allowed (Marcher::ray,
// src/calc.rs:2328-2355), Common optic: pixel_ray at non-integer pixels.
__device__ __forceinline__ void frac_ray(const DevP &P, double fx, double fy, double c[3], double v[3]) {
    const double cafy = (fy / (double)P.height - 0.5) * P.fov_y;
    const double cafx = (double)((P.fovx_off - (float)(fx + 1.0)) * P.fovx_mul);
    double sx, cx, sy, cy;
    sincos(cafy, &sx, &cx);
    sincos(cafx, &sy, &cy);
    double w0 = -sy, w1 = sx, w2 = cx * cy;
    const double nrm = 1.0 / sqrt(w0 * w0 + w1 * w1 + w2 * w2);
    w0 *= nrm;
    w1 *= nrm;
    w2 *= nrm;
    for (int k = 0; k < 3; k++) v[k] = w0 * P.vg[0][k] + w1 * P.vg[1][k] + w2 * P.vg[2][k];
    for (int k = 0; k < 3; k++) c[k] = P.ystart[k] + P.vg[0][k] * fx + P.vg[1][k] * fy;
}

// depth from a G-buffer depth code at the given fraction of its last digit
// (src/calc.rs:2167, 2748-2750)
__device__ __forceinline__ double depth_of_code(const DevP &P, double zf, double frac) {
    const double a = (frac - zf) / P.zc_mul + 1.0;
    return (a * a - 1.0) / P.zcorr;
}

}  // namespace mb3d
