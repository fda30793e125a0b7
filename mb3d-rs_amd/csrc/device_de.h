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
// (src/formulas.rs:707-736).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/mb3d_hip.h"

namespace mb3d {

enum Kind : int { K_P8 = 0, K_BOX = 1, K_HYB = 2 };
enum DeMode : int { DEM_NUMERIC = 0, DEM_BOX = 1, DEM_P8 = 2 };

// The march parameters as the kernels take them (by value, so that the
// uniform ones live in scalar registers).
struct DevP {
    int width, height, row0, rows;
    int n_slots, end_to, repeat_from, max_it, min_it;
    int de_option, is_custom_de, de_add_steps, dfog_on_it, first_step_random, step_cap;
    int tile_w, tile_h, tiles_x;
    int slot_kind[MB3D_HIP_MAX_SLOTS], slot_iters[MB3D_HIP_MAX_SLOTS], slot_ade[MB3D_HIP_MAX_SLOTS];
    double slot_z_mul[MB3D_HIP_MAX_SLOTS], slot_scale[MB3D_HIP_MAX_SLOTS];
    double slot_min_r[MB3D_HIP_MAX_SLOTS], slot_fold[MB3D_HIP_MAX_SLOTS];
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
};

// Per-lane work counters, reduced per wave at the end of a kernel.
struct Cnt {
    uint32_t steps = 0, de32 = 0, de64 = 0;
    uint32_t p8_32 = 0, box_32 = 0, p8_64 = 0, box_64 = 0;
    uint32_t hits = 0, overflow = 0;
};

// ---------------------------------------------------------------------------
// FP32 formula iterations

__device__ __forceinline__ void p8f(float &x, float &y, float &z, float jx, float jy, float jz, float zmul8) {
#pragma clang fp contract(fast)
    const float xx = x * x, yy = y * y, zz = z * z;
    const float r = xx + yy, rr = r * r, zzzz = zz * zz;
    const float t = (zzzz - 6.0f * r * zz + rr) * (zz - r) * __builtin_amdgcn_sqrtf(r);
    const float nz = jz - t * z * zmul8;
    const float num = (rr * 70.0f + zzzz) * zzzz - 28.0f * zz * r * (zzzz + rr);
    const float a = num * __builtin_amdgcn_rcpf(rr * rr + 1e-37f) + 1.0f;
    const float xxxx = xx * xx, yyyy = yy * yy;
    const float ny = (yyyy * (7.0f * xx - yy) + xxxx * (xx - 7.0f * yy)) * 8.0f * x * y * a + jy;
    const float nx = a * (xxxx * xxxx + yyyy * (70.0f * xxxx + yyyy) - 28.0f * (xx * yy * (yyyy + xxxx))) + jx;
    x = nx;
    y = ny;
    z = nz;
}

__device__ __forceinline__ float bfoldf(float x, float f) { return fabsf(x + f) - (fabsf(x - f) + x); }

// amazing_box in FP32; returns the derivative factor
__device__ __forceinline__ float boxf(float &x, float &y, float &z, float jx, float jy, float jz, float fold,
                                     float scale, float min_r2, float scale_min_r2) {
#pragma clang fp contract(fast)
    const float a = bfoldf(x, fold), b = bfoldf(y, fold), c = bfoldf(z, fold);
    const float r = c * c + b * b + a * a;
    const float mul = r < min_r2 ? scale_min_r2 : (1.0f < r ? scale : scale * __builtin_amdgcn_rcpf(r));
    x = a * mul + jx;
    y = b * mul + jy;
    z = c * mul + jz;
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

// The numeric gradient DE in FP32 (calc_de_part, src/calc.rs:1515-1545):
// the point and its three offset points iterated in one loop.
template <int KIND>
__device__ float de32_numeric(const DevP &P, const double c[3], float ms, int &it_out, Cnt &cnt) {
#pragma clang fp contract(fast)
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
            for (int k = 0; k < 4; k++) p8f(x[k], y[k], z[k], jx[k], jy[k], jz[k], zm);
            cnt.p8_32 += 1u + a1 + a2 + a3;
        } else {
            const float fo = P.ffold[slot], sc = P.fscale[slot], m2 = P.fmin_r2[slot], sm = P.fscale_min_r2[slot];
#pragma unroll
            for (int k = 0; k < 4; k++) boxf(x[k], y[k], z[k], jx[k], jy[k], jz[k], fo, sc, m2, sm);
            cnt.box_32 += 1u + a1 + a2 + a3;
        }
        it++;
        r0 = x[0] * x[0] + y[0] * y[0] + z[0] * z[0];
        const float q1 = x[1] * x[1] + y[1] * y[1] + z[1] * z[1];
        const float q2 = x[2] * x[2] + y[2] * y[2] + z[2] * z[2];
        const float q3 = x[3] * x[3] + y[3] * y[3] + z[3] * z[3];
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
        const float g = (r0 - r1) * (r0 - r1) + (r0 - r2) * (r0 - r2) + (r0 - r3) * (r0 - r3);
        res = r0 * __logf(r0) * P.d_de_scale / (__builtin_amdgcn_sqrtf(g) + P.de_offset006);
    }
    const float mn = ms * 0.25f;
    if (!(res >= mn)) res = mn;  // also NaN
    return res;
}

// The analytic DEs in FP32: Amazing Box |z| / |dr| (option 11,
// src/iteration.rs:562-568) and the power-8 extension 0.5 r ln r / dr.
template <int DEM>
__device__ float de32_analytic(const DevP &P, const double c[3], float ms, int &it_out, Cnt &cnt) {
#pragma clang fp contract(fast)
    const float jx = (float)c[0], jy = (float)c[1], jz = (float)c[2];
    float x = jx, y = jy, z = jz;
    float r = x * x + y * y + z * z;
    float w = 1.0f;
    int it = 0;
    if (DEM == DEM_BOX) {
        const float fo = P.ffold[0], sc = P.fscale[0], m2 = P.fmin_r2[0], sm = P.fscale_min_r2[0];
        for (;;) {
            w *= boxf(x, y, z, jx, jy, jz, fo, sc, m2, sm);
            it++;
            r = x * x + y * y + z * z;
            if (it >= P.max_it || r > P.f_d_rstop) break;
        }
        cnt.box_32 += it;
    } else {
        const float zm8 = 8.0f * P.fz_mul[0];
        for (;;) {
            const float r2 = x * x + y * y + z * z;
            w = 8.0f * (r2 * r2 * r2 * __builtin_amdgcn_sqrtf(r2)) * w + 1.0f;
            p8f(x, y, z, jx, jy, jz, zm8);
            it++;
            r = x * x + y * y + z * z;
            if (it >= P.max_it || r > P.f_d_rstop) break;
        }
        cnt.p8_32 += it;
    }
    it_out = it;
    cnt.de32++;
    float res;
    if (DEM == DEM_BOX)
        res = __builtin_amdgcn_sqrtf(r) * __builtin_amdgcn_rcpf(fabsf(w)) * P.d_de_scale;
    else
        res = __builtin_amdgcn_sqrtf(r) * 0.25f * __logf(r) * __builtin_amdgcn_rcpf(w) * P.d_de_scale;
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

struct M64 {
    // Iteration (src/iteration.rs:59-105), the parts the built-ins use
    double v[4], c[3], j[3];
    double rout, rold, otrap, deriv1;
    int it_result, max_it;
    double rstop;
    bool calc_sit;
    float smooth_it;
    // Marcher (src/calc.rs:1143-1170)
    double mzz, vfov[3];
    float ms_de_stop;
    int max_its_result;

    __device__ void init(const DevP &P) {
        v[0] = v[1] = v[2] = v[3] = 0.0;
        rout = rold = otrap = 0.0;
        deriv1 = 1.0;
        it_result = 0;
        max_it = P.max_it;
        rstop = P.d_rstop;
        calc_sit = false;
        smooth_it = 0.0f;
        mzz = 0.0;
        ms_de_stop = P.de_stop;
        max_its_result = P.max_it;
    }

    // Formula::iterate via call_slot (src/iteration.rs:325-345)
    __device__ void call_slot(const DevP &P, int n, Cnt &cnt) {
        if (P.slot_kind[n] == MB3D_HIP_FORMULA_INTPOW) {
            if (P.slot_ade[n]) {  // the analytic power-8 extension: dr <- 8 r^7 dr + 1
                const double r2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
                deriv1 = 8.0 * (r2 * r2 * r2 * sqrt(r2)) * deriv1 + 1.0;
            }
            const double x = v[0], y = v[1], z = v[2];
            const double xx = x * x, yy = y * y, zz = z * z;
            const double r = xx + yy;
            const double rr = r * r;
            const double zzzz = zz * zz;
            const double t = (zzzz - 6.0 * r * zz + rr) * (zz - r) * sqrt(r);
            v[2] = -(t * z * 8.0 * P.slot_z_mul[n]) + j[2];
            const double num = (rr * 70.0 + zzzz) * zzzz - 28.0 * zz * r * (zzzz + rr);
            const double a = num / (rr * rr + 1e-40) + 1.0;
            const double xxxx = xx * xx, yyyy = yy * yy;
            v[1] = (yyyy * (7.0 * xx - yy) + xxxx * (xx - 7.0 * yy)) * 8.0 * x * y * a + j[1];
            v[0] = a * (xxxx * xxxx + yyyy * (70.0 * xxxx + yyyy) - 28.0 * (xx * yy * (yyyy + xxxx))) + j[0];
            cnt.p8_64++;
        } else {
            const double fold = P.slot_fold[n], scale = P.slot_scale[n];
            const double x = bfold64(v[0], fold), y = bfold64(v[1], fold), z = bfold64(v[2], fold);
            const double r = z * z + y * y + x * x;
            const double m = fmax(P.slot_min_r[n], 1e-40);
            const double sqr_min_r = m * m;
            double mul;
            if (r < sqr_min_r)
                mul = scale / sqr_min_r;
            else if (1.0 < r)
                mul = scale;
            else
                mul = scale / r;
            if (P.slot_ade[n]) v[3] *= mul;
            v[2] = z * mul + j[2];
            v[1] = y * mul + j[1];
            v[0] = x * mul + j[0];
            cnt.box_64++;
        }
    }

    // calc_smooth_iterations (src/iteration.rs:140-154)
    __device__ void smooth(const DevP &P, int n) {
        if (rout <= 1.0) {
            smooth_it = (float)it_result;
        } else if (rold < 1.0) {
            const double d = log(log(rout) * 0.5) * (double)P.fhln[n];
            smooth_it = (float)((double)it_result + (double)P.ln_rstop - d);
        } else {
            const double d = log(0.5 * log(rout));
            const double e = d - log(0.5 * log(rold));
            smooth_it = (float)(((double)P.ln_rstop - d) / (e + 1e-100) + (double)it_result);
        }
    }

    // run_with (src/iteration.rs:268-315), 3D, counted slots
    __device__ int run(const DevP &P, Cnt &cnt) {
        int n = 0;
        int btmp = P.slot_iters[0];
        it_result = 0;
        for (;;) {
            rold = rout;
            int guard = 0;
            while (btmp <= 0) {
                n += 1;
                if (n > P.end_to) n = P.repeat_from;
                btmp = P.slot_iters[n];
                if (++guard > 64) return n;
            }
            call_slot(P, n, cnt);
            btmp -= 1;
            it_result += 1;
            rout = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
            if (rout < otrap) otrap = rout;
            if (it_result >= max_it || rout > rstop) return n;
        }
    }

    __device__ void start_point() {
        j[0] = c[0];
        j[1] = c[1];
        j[2] = c[2];
        v[0] = c[0];
        v[1] = c[1];
        v[2] = c[2];
        rout = c[0] * c[0] + c[1] * c[1] + c[2] * c[2];
        otrap = rout;
    }

    // hybrid_3d (src/iteration.rs:521-537)
    __device__ void hybrid_3d(const DevP &P, Cnt &cnt) {
        start_point();
        v[3] = 0.0;
        const int n = run(P, cnt);
        if (calc_sit) smooth(P, n);
    }

    // hybrid_3d_de (src/iteration.rs:539-585); option 32: the textbook
    // 0.5 r ln r / dr of the power-8 extension
    __device__ double hybrid_3d_de(const DevP &P, Cnt &cnt) {
        start_point();
        if ((P.de_option & 0x38) == 32)
            deriv1 = 1.0;
        else if ((P.de_option & 0x38) == 16)
            v[3] = rout;
        else
            v[3] = 1.0;
        const int n = run(P, cnt);
        double result;
        if ((P.de_option & 0x38) == 32)
            result = sqrt(rout) * 0.25 * log(rout) / deriv1;
        else
            result = sqrt(rout) / fabs(v[3]);
        if (calc_sit) smooth(P, n);
        return result;
    }

    // calc_de_part without dIFS (src/calc.rs:1495-1559)
    __device__ __noinline__ double calc_de(const DevP &P, Cnt &cnt) {
        cnt.de64++;
        double result;
        if (P.is_custom_de) {
            result = hybrid_3d_de(P, cnt) * (double)P.d_de_scale;
        } else {
            hybrid_3d(P, cnt);
            if (rout < 1e-200) {
                result = 0.0;
            } else {
                const bool buf_sit = calc_sit;
                const int buf_max_it = max_it;
                const double buf_rout = rout;
                calc_sit = false;
                max_it = it_result;
                rstop = P.rstop3d;
                const double off = (double)P.de_offset;
                double g[3];
                for (int k = 0; k < 3; k++) {
                    const double bufd = c[k];
                    c[k] = bufd + off;
                    hybrid_3d(P, cnt);
                    g[k] = (buf_rout - rout) * (buf_rout - rout);
                    c[k] = bufd;
                }
                result = buf_rout * log(buf_rout) * (double)P.d_de_scale /
                         (sqrt(g[0] + g[1] + g[2]) + (double)P.de_offset006);
                rout = buf_rout;
                it_result = max_it;
                max_it = buf_max_it;
                calc_sit = buf_sit;
                rstop = P.d_rstop;
            }
        }
        max_its_result = max_it;
        const double mn = (double)ms_de_stop * 0.25;
        if (result < mn || isnan(result)) result = mn;
        return result;
    }

    __device__ __forceinline__ void step(double d) {
        c[0] += vfov[0] * d;
        c[1] += vfov[1] * d;
        c[2] += vfov[2] * d;
    }

    // update_de_stop (src/calc.rs:1698-1703)
    __device__ __forceinline__ void update_de_stop(const DevP &P) {
        ms_de_stop = (float)((double)P.de_stop * (1.0 + mzz * (double)P.de_stop_factor));
    }

    // bin_search (src/calc.rs:1706-1735)
    __device__ void bin_search(const DevP &P, double &de, double last_step_width, Cnt &cnt) {
        int itmp = P.de_add_steps;
        double dt1 = last_step_width * -0.5;
        while (fabs(de - (double)ms_de_stop) > 0.001) {
            mzz += dt1;
            step(dt1);
            update_de_stop(P);
            itmp -= 1;
            if (itmp <= 0) break;
            de = calc_de(P, cnt);
            if (it_result >= max_its_result) {
                dt1 = -fabs(dt1);
                continue;
            }
            const double f = de - (double)ms_de_stop;
            dt1 = f < 0.0 ? fabs(dt1) * -0.55 : fabs(dt1) * 0.55;
        }
    }

    // bin_search_it (src/calc.rs:1737-1792)
    __device__ __noinline__ void bin_search_it(const DevP &P, int steps, Cnt &cnt) {
        const float yp = (float)max_it - 0.99f;
        const int saved_max = max_it;
        max_it += 1;
        int itmp = steps;
        calc_sit = true;
        double dt1 = 0.0, dmul = 1.0;
        bool first = true;
        float last_si = 0.0f, last_dif = 0.0f;
        for (;;) {
            mzz += dt1;
            step(dt1);
            calc_de(P, cnt);
            if (!first && last_dif < fabsf(yp - smooth_it)) {
                mzz -= dt1;
                step(-dt1);
                smooth_it = last_si;
                dmul *= dt1 > 0.0 ? 0.5 : 0.7;
            }
            last_dif = fabsf(yp - smooth_it);
            last_si = smooth_it;
            if (smooth_it > (float)max_it - 0.1f) {
                dt1 = -3.0;
            } else {
                mzz -= 0.001;
                step(-0.001);
                calc_de(P, cnt);
                const float r = last_si - smooth_it;
                if (fabsf(r) < 1e-30f)
                    dt1 = (last_si < smooth_it ? 1.0 : 0.0) - 0.5;
                else if (r < 0.0f)
                    dt1 = (double)((yp - smooth_it) / (r * 500.0f));
                else
                    dt1 = (double)((yp - smooth_it) / (r * 1000.0f));
                if (dt1 > 4.0)
                    dt1 = sqrt(dt1) * 2.0;
                else if (dt1 < -9.0)
                    dt1 = sqrt(-dt1) * -3.0;
                dt1 = dt1 * dmul + 0.0005;
            }
            first = false;
            itmp -= 1;
            if (itmp < 0) break;
        }
        max_it = saved_max;
    }

    // calculate_normals, smooth_normals = 0 (src/calc.rs:1813-1858)
    __device__ float normals(const DevP &P, double n[3], Cnt &cnt) {
        calc_sit = true;
        const double noffset = (double)fminf(P.de_stop, 1.0f) * (1.0 + mzz * (double)P.de_stop_factor) * 0.15;
        const double ct[3] = {c[0], c[1], c[2]};
        calc_de(P, cnt);
        const float nn = smooth_it;
        calc_sit = false;
#pragma unroll
        for (int q = 0; q < 3; q++) {
            const int axis = q == 0 ? 2 : q - 1;  // the order vg2, vg0, vg1
            c[0] = ct[0] + P.vg[axis][0] * noffset;
            c[1] = ct[1] + P.vg[axis][1] * noffset;
            c[2] = ct[2] + P.vg[axis][2] * noffset;
            const double a = calc_de(P, cnt);
            const double m2 = -2.0 * noffset;
            c[0] += P.vg[axis][0] * m2;
            c[1] += P.vg[axis][1] * m2;
            c[2] += P.vg[axis][2] * m2;
            n[axis] = (a - calc_de(P, cnt)) * 0.5;
        }
        c[0] = ct[0];
        c[1] = ct[1];
        c[2] = ct[2];
        return nn;
    }

    // calc_zpos_and_rough, smooth_normals = 0 (src/calc.rs:2194-2205)
    __device__ uint32_t zpos_fine(const DevP &P) const {
        const double zz = fmax(mzz, 0.0);
        const double vr = rte(P.zc_mul * (sqrt(zz * P.zcorr + 1.0) - 1.0));
        const int64_t vv = (int64_t)vr;
        int64_t itmp = 8388352 - vv;
        itmp = itmp < 0 ? 0 : (itmp > 8388352 ? 8388352 : itmp);
        return (uint32_t)itmp << 8;
    }
};

// The ray of pixel (x, y) (calc_vgrads_fov + calc_start_pos,
// src/calc.rs:1561-1575,1673-1690; build_view_vector_dfov, rotate_vector_reverse
// src/math.rs:63-69,127-133), Common optic.
__device__ __forceinline__ void pixel_ray(const DevP &P, int x, int y, double c[3], double v[3]) {
    const double cafy = ((double)y / (double)P.height - 0.5) * P.fov_y;
    const double cafx = (double)((P.fovx_off - (float)(x + 1)) * P.fovx_mul);
    const double sx = sin(cafy), cx = cos(cafy);
    const double sy = sin(cafx), cy = cos(cafx);
    double w0 = -sy, w1 = sx, w2 = cx * cy;
    const double nrm = 1.0 / sqrt(w0 * w0 + w1 * w1 + w2 * w2);
    w0 *= nrm;
    w1 *= nrm;
    w2 *= nrm;
    for (int k = 0; k < 3; k++) v[k] = w0 * P.vg[0][k] + w1 * P.vg[1][k] + w2 * P.vg[2][k];
    const double fx = x, fy = y;
    for (int k = 0; k < 3; k++) c[k] = P.ystart[k] + P.vg[0][k] * fx + P.vg[1][k] * fy;
}

// depth from a G-buffer depth code at the given fraction of its last digit
// (src/calc.rs:2167, 2748-2750)
__device__ __forceinline__ double depth_of_code(const DevP &P, double zf, double frac) {
    const double a = (frac - zf) / P.zc_mul + 1.0;
    return (a * a - 1.0) / P.zcorr;
}

}  // namespace mb3d
