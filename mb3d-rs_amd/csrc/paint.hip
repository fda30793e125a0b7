// paint.hip -- the painter: every G-buffer record of a rectangle lit and
// coloured into RGB (LightVals::pixel_color_s, src/lighting.rs:1739-1760,
// with plv_for_pixel 1762-1798 and shade 1800-2150), one lane per pixel.
//
// The reference paints on the CPU in f32; this kernel keeps its f32
// operations in the same order (no contraction: -ffp-contract=off), so the
// colours come out the same but for the f32 sine, cosine and arcsine of
// the view rays (device: the f64 functions rounded to f32).
//
// Layout: a block of 256 lanes takes 256 consecutive pixels per turn.  The
// 20-byte records are staged through LDS with unit-stride dword loads, and
// the block's colours (3, 6 or 12 bytes a pixel) leave as dwords.  The
// light values and the 8 x 128 diffuse tables sit in LDS for the block's
// lifetime.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>

#include "../../include/mb3d_hip.h"
#include "host_common.h"
#include "paint_common.h"

namespace mb3d {

namespace {

constexpr int kPaintBlock = 256;

__device__ float g_cos_tab[8 * 128];

struct Paint {
    const mb3d_hip_light_vals *lv;
    const mb3d_hip_paint_camera *cam;
    const float *tab;  // [8][128]
};

// ---- f32 helpers with the reference's semantics ----

// Rust's f32::clamp (a NaN passes through)
__device__ __forceinline__ float clamp_rs(float v, float lo, float hi) {
    if (v < lo) v = lo;
    if (v > hi) v = hi;
    return v;
}

// f32 sine, cosine, arcsine: the f64 function rounded to f32 (glibc's f32
// functions, which the reference calls, round the exact value the same
// way but in rare ties of the last bit)
__device__ __forceinline__ float sin_f(float x) { return (float)sin((double)x); }
__device__ __forceinline__ float cos_f(float x) { return (float)cos((double)x); }
__device__ __forceinline__ float asin_f(float x) { return (float)asin((double)x); }

// `x.powi(n)`: multiplications by squaring from the low bit up (LLVM's
// expansion of powi and compiler-rt's __powisf2)
__device__ __forceinline__ float powi_f(float x, int n) {
    float res = 0.0f, sq = x;
    bool have = false;
    while (n) {
        if (n & 1) {
            res = have ? res * sq : sq;
            have = true;
        }
        n >>= 1;
        if (n) sq = sq * sq;
    }
    return res;
}

// Rust's saturating `as i32` of an f32
__device__ __forceinline__ int to_i32(float v) {
    if (v != v) return 0;
    if (v >= 2147483648.0f) return 2147483647;
    if (v <= -2147483648.0f) return (int)0x80000000u;
    return (int)v;
}

// FastIntPow (src/math.rs:273-285): base^expo for expo a power of two
__device__ __forceinline__ float fast_int_pow(float base, int expo) {
    if (base < 0.0f) return 0.0f;
    float r = base;
    for (int i = expo >> 1; i > 0; i >>= 1) r *= r;
    return r;
}

// GetCosTabVal / GetCosTabValSqr (src/lighting.rs:511-541) with
// MakeSplineCoeff (src/math.rs:263-270)
__device__ float cos_tab_val(const float *tab, int tnr, float dotp, float rough, bool sqr) {
    float t = 62.0f - 60.0f * dotp;
    int ip = to_i32(truncf(t)) - 1;
    if (ip < 0) {
        ip = 0;
        t = 0.0f;
    } else if (ip > 124) {
        ip = 124;
        t = 1.0f;
    } else {
        t = t - truncf(t);
    }
    const double xs = (double)t;
    const double r3 = xs * xs * xs / 6.0;
    const double r0 = 1.0 / 6.0 + 0.5 * xs * (xs - 1.0) - r3;
    const double r2 = xs + r0 - 2.0 * r3;
    const double r1 = 1.0 - r0 - r2 - r3;
    const float w[4] = {(float)r0, (float)r1, (float)r2, (float)r3};
    const float *p1 = tab + tnr * 128 + ip;
    const float *p2 = tab + (tnr + 4) * 128 + ip;
    float a = p1[0] * w[0], b = p2[0] * w[0];
    for (int k = 1; k < 4; k++) {
        a = a + p1[k] * w[k];
        b = b + p2[k] * w[k];
    }
    if (sqr) {
        a *= a;
        b *= b;
    }
    return a + rough * (b - a);
}

// CalcPosLightShape (pos_light_shape, src/lighting.rs:899-941)
__device__ void pos_light_shape(float &flux, float &transp, int func, bool pos, bool sqr) {
    const float tmp_r = transp;
    const float rtmp_r = 1.0f / tmp_r;
    if (func == 4) {
        const float t = flux * rtmp_r;
        transp = (t * t) * (t * t);
        flux = 7e-3f * (tmp_r - flux) * rtmp_r * rtmp_r;
    } else if (func == 2) {
        transp = 1.0f;
        const float f = 1.0f - sqrtf(flux * rtmp_r);
        flux = 5e-2f * powi_f(fmaxf(f, 0.0f), 50) * rtmp_r;
        if (pos) flux *= 1.5f;
    } else if (func == 6) {
        const float f = flux * 1.05f * rtmp_r;
        if (f < 1.0f) {
            transp = pos ? fmaxf(f - 0.95f, 0.0f) * 10.0f : 0.7f;
            flux = 4e-3f * sqrtf(1.05f - f * f) * rtmp_r;
            if (!pos) flux *= 0.7f;
        } else if (pos) {
            transp = (f - 0.95f) * 10.0f;
            flux = 17.8e-3f * (1.05f - f) * rtmp_r;
        } else {
            transp = (f - 1.0f) * 6.0f + 0.7f;
            flux = 12.5e-3f * (1.05f - f) * rtmp_r;
        }
    } else {
        transp = 1.0f;
        const float t = flux * rtmp_r;
        const float k = sqr ? 40.0f : 25.0f;
        const float q = (t + 0.0004f) * k;
        flux = (1.0f - t) * 0.1f / (q * q) * rtmp_r;
    }
}

// ConvertVLight (src/lighting.rs:944-948)
__device__ __forceinline__ float convert_vlight(uint32_t w) {
    w &= 0x3FFu;
    return (float)((w & 0x7Fu) << ((w >> 7) & 7u));
}

__device__ __forceinline__ float dot3(const float a[3], const float b[3]) {
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

// sv2 + w1 (sv1 - sv2) (LinInterpolate2SVecs)
__device__ __forceinline__ void lerp3(const float sv1[3], const float sv2[3], float w1, float o[3]) {
    for (int k = 0; k < 3; k++) o[k] = sv2[k] + w1 * (sv1[k] - sv2[k]);
}

__device__ __forceinline__ float luma(const float v[3]) { return v[0] * 0.3f + v[1] * 0.59f + v[2] * 0.11f; }

// PaintCamera::view_vec (src/lighting.rs:748-772)
__device__ void view_vec(const mb3d_hip_paint_camera &c, float x_pos, float y_pos, float aspect, float v[3]) {
    const float cx = (c.x_off - x_pos) * c.fov * aspect;
    const float cy = (y_pos - 0.5f) * c.fov;
    if (c.planar == 1) {
        const float t[3] = {-cx, cy, c.pl_optic_z};
        const float n = sqrtf(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
        const float r = 1.0f / n;
        for (int k = 0; k < 3; k++) v[k] = t[k] * r;
    } else if (c.planar == 2) {
        v[0] = -sin_f(cx) * cos_f(cy);
        v[1] = sin_f(cy);
        v[2] = cos_f(cy) * cos_f(cx);
    } else if (c.planar == 3) {
        const double xa = cy, ya = cx;
        const double t = sqrt(xa * xa + ya * ya);
        if (t < 1e-12) {
            v[0] = 0.0f, v[1] = 0.0f, v[2] = 1.0f;
        } else {
            const double tc = fmin(t, 3.141592653589793);
            const double s = sin(tc) / t;
            v[0] = (float)(-ya * s), v[1] = (float)(xa * s), v[2] = (float)cos(tc);
        }
    } else {
        const float t[3] = {-sin_f(cx), sin_f(cy), cos_f(cy) * cos_f(cx)};
        const float n = sqrtf(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
        const float r = 1.0f / n;
        for (int k = 0; k < 3; k++) v[k] = t[k] * r;
    }
}

__device__ __forceinline__ void to_abs(const mb3d_hip_paint_camera &c, const float v[3], float o[3]) {
    for (int k = 0; k < 3; k++) o[k] = v[0] * c.m[0][k] + v[1] * c.m[1][k] + v[2] * c.m[2][k];
}

__device__ __forceinline__ void to_view(const mb3d_hip_paint_camera &c, const float v[3], float o[3]) {
    for (int k = 0; k < 3; k++) o[k] = dot3(c.m[k], v);
}

__device__ __forceinline__ void amb_abs(const mb3d_hip_light_vals &lv, const float v[3], float o[3]) {
    if (lv.has_amb_frame) {
        for (int k = 0; k < 3; k++) o[k] = dot3(lv.amb_frame[k], v);
    } else {
        for (int k = 0; k < 3; k++) o[k] = v[k];
    }
}

// The pixel's G-buffer words (the card record, src/gpu.rs:3452-3463)
struct Si {
    float n[3];
    uint32_t zpos_fine, shadow, amb_shadow, si_gradient, otrap;
};

// CalcColors (calc_colors_alpha + palette_place, src/lighting.rs:1560-1628):
// diffuse and specular colour
__device__ void calc_colors(const mb3d_hip_light_vals &lv, const Si &si, float idif0, float dif[3], float spe[3]) {
    if (si.otrap == 0xFFFFu && lv.has_ground) {
        for (int k = 0; k < 3; k++) dif[k] = lv.ground_dif[k], spe[k] = lv.ground_spe[k];
        return;
    }
    const float irf = lv.col_on_otrap ? (float)(si.otrap & 0x7FFFu) : (float)si.si_gradient;
    int ir = (int)roundf(clamp_rs(((irf - lv.s_c_start) * lv.s_c_mul + idif0) * 16384.0f, -1e9f, 1e9f));
    int single = -1;
    if (lv.col_cycling)
        ir &= 32767;
    else if (ir < 0)
        single = 0;
    else if (ir >= lv.col_pos[9])
        single = 9;
    int il1 = 0, il2 = 0;
    float t = 0.0f;
    if (single < 0) {
        il2 = 5;
        if (lv.col_pos[il2] < ir) {
            do {
                il2++;
            } while (il2 != 10 && lv.col_pos[il2] < ir);
        } else {
            while (il2 > 1 && lv.col_pos[il2 - 1] >= ir) il2--;
        }
        if (lv.no_col_ipol) {
            single = il2 - 1;
        } else {
            il1 = il2 - 1;
            if (il2 > 9) il2 = 0;
            t = (float)(ir - lv.col_pos[il1]) * lv.s_c_div[il1];
        }
    }
    if (single >= 0) {
        for (int k = 0; k < 3; k++) dif[k] = lv.col_dif[single][k], spe[k] = lv.col_spe[single][k];
        return;
    }
    lerp3(lv.col_dif[il2], lv.col_dif[il1], t, dif);
    lerp3(lv.col_spe[il2], lv.col_spe[il1], t, spe);
}

// CalcColorsInside (src/lighting.rs:1675-1705): diffuse colour, specular amount
__device__ void calc_colors_inside(const mb3d_hip_light_vals &lv, const Si &si, float idif0, float dif[3],
                                   float spe[3]) {
    int ir = to_i32(roundf((((float)si.si_gradient - lv.s_ci_start) * lv.s_ci_mul + idif0) * 16384.0f));
    float sp;
    const bool cyc = lv.col_cycling != 0;
    if (cyc) ir &= 32767;
    if (!cyc && ir < 0) {
        for (int k = 0; k < 3; k++) dif[k] = lv.col_int[0][k];
        sp = lv.col_int_spec[0];
    } else if (!cyc && ir > lv.icol_pos[3]) {
        for (int k = 0; k < 3; k++) dif[k] = lv.col_int[3][k];
        sp = lv.col_int_spec[3];
    } else {
        int il2 = 1;
        while (il2 < 4 && lv.icol_pos[il2] < ir) il2++;
        if (lv.no_col_ipol) {
            for (int k = 0; k < 3; k++) dif[k] = lv.col_int[il2 - 1][k];
            sp = lv.col_int_spec[il2 - 1];
        } else {
            const int il1 = il2 - 1;
            il2 &= 3;
            const float t = (float)(ir - lv.icol_pos[il1]) * lv.s_ic_div[il1];
            lerp3(lv.col_int[il2], lv.col_int[il1], t, dif);
            sp = lv.col_int_spec[il2] + t * (lv.col_int_spec[il1] - lv.col_int_spec[il2]);
        }
    }
    spe[0] = spe[1] = spe[2] = sp;
}

// plv_for_pixel + shade (own pixel, no reflection options) for one record:
// the light on the 0..255 scale before clipping and gamma
__device__ void shade(const Paint &P, const Si &si, int x, int y, float out[3]) {
    const mb3d_hip_light_vals &lv = *P.lv;
    const mb3d_hip_paint_camera &cam = *P.cam;
    // ---- plv_for_pixel
    const float x_pos = (float)(x + 1) / (float)cam.width;
    const float y_pos = (float)y / (float)cam.height;
    float view[3];
    view_vec(cam, x_pos, y_pos, cam.aspect, view);
    const float ys = lv.depth_func == 1 ? y_pos * y_pos : lv.depth_func == 0 ? y_pos : sqrtf(fmaxf(y_pos, 0.0f));
    const uint32_t zpos_word = si.zpos_fine >> 16;
    double z1;
    if (zpos_word > 32767u) {
        const double q = 8388352.0 / cam.zc_mul + 1.0;
        z1 = (q * q - 1.0) * cam.step_width / cam.zcorr;
    } else {
        const double q = (double)(8388352ll - (long long)(si.zpos_fine >> 8)) / cam.zc_mul + 1.0;
        z1 = (q * q - 1.0) * cam.step_width / cam.zcorr;
    }
    const float z_pos = (float)(z1 + cam.zz_stmit_dif);
    float abs_view[3], cam_pos[3], obj_pos[3], dep_c[3];
    to_abs(cam, view, abs_view);
    if (cam.planar == 2) {
        for (int k = 0; k < 3; k++) cam_pos[k] = cam.start_pos[k];
    } else {
        for (int k = 0; k < 3; k++)
            cam_pos[k] = (cam.start_pos[k] + cam.y_add[k] * (float)y) + cam.x_add[k] * (float)x;
    }
    const float z1f = (float)z1;
    for (int k = 0; k < 3; k++) obj_pos[k] = cam_pos[k] + abs_view[k] * z1f;
    lerp3(lv.depth_col2, lv.depth_col, ys, dep_c);

    // ---- shade
    if (lv.amb_rel_obj) {
        float aa[3];
        amb_abs(lv, abs_view, aa);
        const float a = asin_f(clamp_rs(aa[1], -1.0f, 1.0f));
        float w;
        if (lv.amb_fov > 0.0f) {
            const float yy = clamp_rs(a / lv.amb_fov + 0.5f, 0.0f, 1.0f);
            w = lv.depth_func == 1 ? yy * yy : lv.depth_func == 0 ? yy : sqrtf(yy);
        } else {
            w = a * 0.31830987f + 0.5f;
        }
        lerp3(lv.depth_col2, lv.depth_col, w, dep_c);
    }
    const bool sqr = lv.sqr != 0;
    float o[3];
    float dtmp;
    if (zpos_word < 32768u) {
        const float d_rough = (float)(si.zpos_fine & 0xFFu) * lv.s_roughness_factor;
        const float shadow = si.amb_shadow >= 16383u ? 1.0f : (float)si.amb_shadow / 16383.0f;
        float d_amb_sh;
        if (lv.s_amb_shad < 1.0f) {
            d_amb_sh = 1.0f - lv.s_amb_shad * shadow;
        } else {
            const float a = 1.0f - shadow;
            d_amb_sh = a + (lv.s_amb_shad - 1.0f) * (a * a - a);
        }
        if (sqr) d_amb_sh = (d_amb_sh * d_amb_sh + d_amb_sh) * 0.5f;
        const float *lns = si.n;
        float li_dif[3] = {0.0f, 0.0f, 0.0f}, li_spe[3] = {0.0f, 0.0f, 0.0f}, li_amb[3] = {0.0f, 0.0f, 0.0f};
        const float d_fog = lv.s_diffuse_shadowing * (d_amb_sh - 1.0f) + 1.0f;
        for (int i = 0; i < lv.n_lights; i++) {
            const mb3d_hip_light_val &L = lv.lights[i];
            const bool soft = L.hs_mask == 0xFFFF;
            const bool no_hs = soft || (si.shadow & (uint32_t)L.hs_mask) == 0u || !L.hs_calced;
            const float soft_mul = soft ? (float)(si.shadow >> 10) * (1.0f / 63.0f) : 1.0f;
            float ln[3];
            float att = 1.0f;
            if (L.positional) {
                float v[3];
                for (int k = 0; k < 3; k++) v[k] = L.ln[k] - obj_pos[k];
                const float d2 = dot3(v, v);
                if (d2 > L.lmax_l) continue;
                const float r = 1.0f / sqrtf(d2);
                for (int k = 0; k < 3; k++) v[k] = v[k] * r;
                to_view(cam, v, ln);
                att = 1.0f / (d2 + 1e-30f);
            } else {
                for (int k = 0; k < 3; k++) ln[k] = L.ln[k];
            }
            const float dotp = lns[0] * ln[0] + lns[1] * ln[1] + lns[2] * ln[2];
            float d = cos_tab_val(P.tab, L.diff_func, dotp, d_rough, sqr) * att;
            for (int k = 0; k < 3; k++) li_amb[k] += L.col[k] * d;
            const float shade_mul = L.sub_amb_sh ? d_amb_sh : d_fog;
            if (no_hs) {
                d *= shade_mul;
                d *= soft_mul;
                for (int k = 0; k < 3; k++) li_dif[k] += L.col[k] * d;
            }
            const float d2 = 2.0f * (lns[0] * view[0] + lns[1] * view[1] + lns[2] * view[2]);
            const float ds = ln[0] * (view[0] - lns[0] * d2) + ln[1] * (view[1] - lns[1] * d2) +
                             ln[2] * (view[2] - lns[2] * d2);
            if (ds > 0.0f) {
                const float pf = (float)L.pow_func;
                float t2 = (att + fminf(d_rough * 2.0f, 1.0f) * (1.0f / pf - att)) * lv.s_spec;
                if (t2 > 0.0f) {
                    const float wa = t2 / pf;
                    for (int k = 0; k < 3; k++) li_amb[k] += L.col[k] * wa;
                    if (no_hs) {
                        t2 *= fast_int_pow(ds, L.pow_func);
                        t2 *= shade_mul;
                        t2 *= soft_mul;
                        for (int k = 0; k < 3; k++) li_spe[k] += L.col[k] * t2;
                    }
                }
            }
        }
        const float idif0 = lv.s_col_zmul * z_pos;
        float idif[3], ispe[3];
        if (si.si_gradient > 32767u && si.otrap != 0xFFFFu)
            calc_colors_inside(lv, si, idif0, idif, ispe);
        else
            calc_colors(lv, si, idif0, idif, ispe);
        // ambient light top / bottom
        float ny;
        if (lv.amb_rel_obj) {
            float a[3], b[3];
            to_abs(cam, lns, a);
            amb_abs(lv, a, b);
            ny = b[1];
        } else {
            ny = lns[1];
        }
        const float t2 = (ny * 0.5f + 0.5f) * d_amb_sh;
        const float t1 = d_amb_sh - t2;
        float iamb[3], dif_s[3], total[3];
        for (int k = 0; k < 3; k++) {
            iamb[k] = (lv.amb_col[k] * t1 + lv.amb_col2[k] * t2) * idif[k];
            dif_s[k] = idif[k] * lv.s_diff;
            total[k] = (iamb[k] + dif_s[k] * li_dif[k]) + ispe[k] * li_spe[k];
        }
        if (!lv.ex_mode) {
            // CalcTotalLight1
            const float s3 = d_amb_sh * (sqr ? 0.11f : 0.3f);
            const float st = fmaxf(1.0f - d_amb_sh, 0.0f);
            const float st2 = st * st * st * st;
            for (int k = 0; k < 3; k++) {
                const float l3 = iamb[k] + (dif_s[k] * li_amb[k]) * s3;
                const float refl = idif[k] * st2 * lv.s_ind_light_reflect + (st - st2);
                o[k] = total[k] + l3 * (idif[k] * refl);
            }
        } else {
            // CalcTotalLight2
            float st = fmaxf(1.0f - d_amb_sh, 0.0f);
            st = st * st * lv.s_ind_light_reflect;
            const float st2 = luma(total) + 4.0f;
            for (int k = 0; k < 3; k++) o[k] = total[k] + (total[k] * idif[k]) * st;
            const float s = (st2 + (st2 + 128.0f) * (st + 0.11f) * 0.04f) / (fmaxf(luma(o), 0.0f) + 3.0f);
            for (int k = 0; k < 3; k++) o[k] = o[k] * s;
        }
        dtmp = fmaxf(((float)zpos_word - 28000.0f) * lv.s_depth + 1.0f, 0.0f);
    } else {
        for (int k = 0; k < 3; k++) o[k] = dep_c[k];
        dtmp = fmaxf(1.0f - (float)(60768 - (int)zpos_word) * lv.s_depth, 0.0f);
    }
    // depth fog and dynamic fog
    if (sqr) {
        if (dtmp < 1.0f) {
            float t = (1.0f - dtmp) * (1.0f - dtmp);
            if (lv.far_fog) t *= t;
            dtmp = 1.0f - t;
        }
    } else if (lv.far_fog && dtmp < 1.0f) {
        dtmp = 1.0f - (1.0f - dtmp) * (1.0f - dtmp);
    }
    for (int k = 0; k < 3; k++) o[k] = o[k] * dtmp;
    const float ir = lv.vol_light ? convert_vlight(si.shadow) : (float)(si.shadow & 0x3FFu);
    float dfog = (ir - lv.s_shad - lv.s_shad_zmul * z_pos) * lv.s_shad_gr;
    if (lv.dfog_options & 2) dfog = fmaxf(dfog, 0.0f);
    float dtmp3 = fminf(ir * lv.s_dyn_fog_mul, 1.0f) * dfog;
    const float fd = fmaxf(1.0f - dtmp, 0.0f);
    for (int k = 0; k < 3; k++) o[k] = o[k] + dep_c[k] * fd;
    if (lv.dfog_options & 1) {
        dfog = clamp_rs(dfog, 0.0f, 1.0f);
        dtmp3 = clamp_rs(dtmp3, 0.0f, 1.0f);
        for (int k = 0; k < 3; k++) o[k] = o[k] * (1.0f - dfog);
    }
    for (int k = 0; k < 3; k++) o[k] = o[k] + (lv.dyn_fog_col[k] * (dfog - dtmp3) + lv.dyn_fog_col2[k] * dtmp3);
    // visible lights
    const int il1 = zpos_word < 32768u ? 32768 - (int)zpos_word : (int)zpos_word;
    const float d_rough_v = 1.0f / (float)il1;
    for (int i = 0; i < lv.n_lights; i++) {
        const mb3d_hip_light_val &L = lv.lights[i];
        if (L.visible == 0) continue;
        const bool pos = L.positional != 0;
        float flux;
        if (pos) {
            float a[3];
            for (int k = 0; k < 3; k++) a[k] = cam_pos[k] - L.ln[k];
            const float c[3] = {a[1] * abs_view[2] - a[2] * abs_view[1], a[2] * abs_view[0] - a[0] * abs_view[2],
                                a[0] * abs_view[1] - a[1] * abs_view[0]};
            flux = dot3(c, c);
        } else {
            if (zpos_word < 32768u) continue;
            flux = fmaxf(1.0f - dot3(L.ln, view), 0.0f);
        }
        float transp = L.lmax_l * 1e-8f;
        if (flux >= transp) continue;
        if (pos) {
            float d[3];
            for (int k = 0; k < 3; k++) d[k] = L.ln[k] - cam_pos[k];
            if (dot3(d, abs_view) < 0.0f) continue;
            if (zpos_word < 32768u) {
                const float t = -sqrtf(transp - flux);
                bool behind;
                if (fabsf(abs_view[0]) > 0.5f)
                    behind = t > (obj_pos[0] - L.ln[0]) / abs_view[0];
                else if (fabsf(abs_view[1]) > 0.5f)
                    behind = t > (obj_pos[1] - L.ln[1]) / abs_view[1];
                else
                    behind = t > (obj_pos[2] - L.ln[2]) / abs_view[2];
                if (behind) continue;
            }
        }
        pos_light_shape(flux, transp, L.visible, pos, sqr);
        float fog_at = 1.0f - fmaxf(1.0f + (L.pos_lp - 28000.0f) * lv.s_depth, 0.0f);
        if (sqr && fog_at > 0.0f) fog_at *= fog_at;
        if (lv.far_fog && fog_at > 0.0f) fog_at *= fog_at;
        flux *= 1.0f - fog_at * 0.9f;
        if (L.visible == 2 || L.visible == 8) {
            for (int k = 0; k < 3; k++) o[k] = o[k] + L.col[k] * flux;
        } else {
            fog_at = fmaxf(fog_at, 0.0f);
            const float il = lv.vol_light ? convert_vlight(si.shadow) : (float)(si.shadow & 0x3FFu);
            const float kk = il * (32768.0f - L.pos_lp) * d_rough_v;
            const float df = (kk - lv.s_shad - lv.s_shad_zmul * (L.pos_z + (float)cam.zz_stmit_dif)) * lv.s_shad_gr *
                             (1.0f - transp) * (1.0f - fog_at);
            const float damb = fminf(kk * lv.s_dyn_fog_mul, 1.0f) * df;
            const float fa = fog_at * (1.0f - transp);
            for (int k = 0; k < 3; k++)
                o[k] = (o[k] * transp + dep_c[k] * fa) +
                       ((lv.dyn_fog_col[k] * (df - damb) + lv.dyn_fog_col2[k] * damb) + L.col[k] * flux);
        }
    }
    for (int k = 0; k < 3; k++) out[k] = o[k];
}

// pixel_color_s: clipping, internal gamma 2, gamma, the channel format
template <int FMT>
__device__ void finish_colour(const mb3d_hip_light_vals &lv, const float o[3], unsigned char *dst) {
    const bool hdr = FMT == MB3D_HIP_PAINT_F32;
    const float top = hdr ? 3.4028234663852886e38f : 255.0f;
    float c[3];
    for (int k = 0; k < 3; k++)
        c[k] = lv.sqr ? sqrtf(clamp_rs(o[k] * 255.0f, 0.0f, top * top)) : clamp_rs(o[k], 0.0f, top);
    if (lv.gamma_h != 0) {
        float g[3];
        for (int k = 0; k < 3; k++) g[k] = lv.gamma_h > 0 ? sqrtf(c[k] * 255.0f) : c[k] * c[k] / 255.0f;
        float r[3];
        lerp3(g, c, lv.s_gamma, r);
        for (int k = 0; k < 3; k++) c[k] = r[k];
    }
    for (int k = 0; k < 3; k++) {
        if (FMT == MB3D_HIP_PAINT_U8) {
            const float v = roundf(c[k]);
            dst[k] = v != v ? 0 : v <= 0.0f ? 0 : v >= 255.0f ? 255 : (unsigned char)v;
        } else if (FMT == MB3D_HIP_PAINT_U16) {
            const float v = roundf(c[k] * 257.0f);
            const unsigned short u = v != v ? 0 : v <= 0.0f ? 0 : v >= 65535.0f ? 65535 : (unsigned short)v;
            std::memcpy(dst + 2 * k, &u, 2);
        } else {
            const float v = c[k] / 255.0f;
            std::memcpy(dst + 4 * k, &v, 4);
        }
    }
}

// Four waves per SIMD (128 VGPRs, a few spills): 64 us against 72 us at the
// compiler's three (163 VGPRs) on a 1080p G-buffer; five and six spill more
// (70 and 68 us)
template <int FMT>
__global__ void __launch_bounds__(kPaintBlock, 4) k_paint(mb3d_hip_paint_camera cam, mb3d_hip_light_vals lv_arg,
                                                       const uint32_t *__restrict__ gbuf, int x0, int y0, int w,
                                                       long long npix, unsigned char *__restrict__ rgb) {
    constexpr int B = FMT == MB3D_HIP_PAINT_U8 ? 3 : FMT == MB3D_HIP_PAINT_U16 ? 6 : 12;  // bytes a pixel
    __shared__ float s_tab[8 * 128];
    __shared__ mb3d_hip_light_vals s_lv;
    __shared__ mb3d_hip_paint_camera s_cam;
    __shared__ uint32_t s_rec[kPaintBlock * 5];
    __shared__ uint32_t s_out[kPaintBlock * B / 4];
    const int tid = threadIdx.x;
    for (int i = tid; i < 8 * 128; i += kPaintBlock) s_tab[i] = g_cos_tab[i];
    {
        const uint32_t *src = reinterpret_cast<const uint32_t *>(&lv_arg);
        uint32_t *dst = reinterpret_cast<uint32_t *>(&s_lv);
        for (int i = tid; i < (int)(sizeof(mb3d_hip_light_vals) / 4); i += kPaintBlock) dst[i] = src[i];
        const uint32_t *csrc = reinterpret_cast<const uint32_t *>(&cam);
        uint32_t *cdst = reinterpret_cast<uint32_t *>(&s_cam);
        for (int i = tid; i < (int)(sizeof(mb3d_hip_paint_camera) / 4); i += kPaintBlock) cdst[i] = csrc[i];
    }
    __syncthreads();
    const Paint P{&s_lv, &s_cam, s_tab};
    const bool aligned = (reinterpret_cast<uintptr_t>(rgb) & 3u) == 0;
    for (long long base = (long long)blockIdx.x * kPaintBlock; base < npix; base += (long long)gridDim.x * kPaintBlock) {
        const int n = (int)min((long long)kPaintBlock, npix - base);
        // the block's records, unit-stride dwords
        const uint32_t *g = gbuf + 5 * base;
        for (int i = tid; i < 5 * n; i += kPaintBlock) s_rec[i] = g[i];
        __syncthreads();
        unsigned char *sb = reinterpret_cast<unsigned char *>(s_out);
        if (tid < n) {
            const uint32_t *r = s_rec + 5 * tid;
            Si si;
            const uint32_t a = r[0], b = r[1], s = r[3];
            si.n[0] = (float)(int16_t)(a & 0xFFFFu) * 3.0518509476e-5f;
            si.n[1] = (float)(int16_t)(a >> 16) * 3.0518509476e-5f;
            si.n[2] = (float)(int16_t)(b & 0xFFFFu) * 3.0518509476e-5f;
            si.shadow = b >> 16;
            si.zpos_fine = r[2];
            si.si_gradient = s & 0xFFFFu;
            si.otrap = s >> 16;
            si.amb_shadow = r[4] & 0xFFFFu;
            const long long i = base + tid;
            const int x = x0 + (int)(i % w), y = y0 + (int)(i / w);
            float o[3];
            shade(P, si, x, y, o);
            finish_colour<FMT>(s_lv, o, sb + B * tid);
        }
        __syncthreads();
        unsigned char *dst = rgb + (size_t)B * base;
        const int bytes = B * n;
        if (aligned) {
            uint32_t *d4 = reinterpret_cast<uint32_t *>(dst);
            for (int i = tid; i < bytes / 4; i += kPaintBlock) d4[i] = s_out[i];
            for (int i = (bytes / 4) * 4 + tid; i < bytes; i += kPaintBlock) dst[i] = sb[i];
        } else {
            for (int i = tid; i < bytes; i += kPaintBlock) dst[i] = sb[i];
        }
        __syncthreads();
    }
}

int g_tab_device = -1;  // the device whose g_cos_tab holds the tables
hipEvent_t g_ev[2] = {nullptr, nullptr};  // the kernel's time, created with the tables

#define PAINT_TRY(expr)                                                                  \
    do {                                                                                 \
        hipError_t e_ = (expr);                                                          \
        if (e_ != hipSuccess) return set_error(MB3D_HIP_ERR_DEVICE, hipGetErrorString(e_)); \
    } while (0)

int paint_locked(const mb3d_hip_paint_camera *cam, const mb3d_hip_light_vals *lv, const uint32_t *gbuf, int x0,
                 int y0, int w, int h, int fmt, void *rgb, int on_device) {
    if (lib_device() < 0) return set_error(MB3D_HIP_ERR_NOT_INIT, "mb3d_hip_init was not called");
    if (!cam || !lv || !gbuf || !rgb) return set_error(MB3D_HIP_ERR_INVALID, "NULL argument");
    if (w < 0 || h < 0 || x0 < 0 || y0 < 0) return set_error(MB3D_HIP_ERR_INVALID, "bad rectangle");
    if (fmt < MB3D_HIP_PAINT_U8 || fmt > MB3D_HIP_PAINT_F32) return set_error(MB3D_HIP_ERR_INVALID, "bad format");
    if (cam->width <= 0 || cam->height <= 0 || cam->planar < 0 || cam->planar > 3)
        return set_error(MB3D_HIP_ERR_INVALID, "bad camera");
    if (lv->n_lights < 0 || lv->n_lights > MB3D_HIP_MAX_LIGHTS) return set_error(MB3D_HIP_ERR_INVALID, "bad light count");
    for (int i = 0; i < lv->n_lights; i++) {
        const mb3d_hip_light_val &L = lv->lights[i];
        if (L.diff_func < 0 || L.diff_func > 3) return set_error(MB3D_HIP_ERR_INVALID, "bad diffuse function");
        if (L.pow_func < 2 || L.pow_func > 256) return set_error(MB3D_HIP_ERR_INVALID, "bad specular exponent");
    }
    const long long npix = (long long)w * h;
    if (npix == 0) return MB3D_HIP_OK;
    hipStream_t st = lib_stream();
    if (g_tab_device != lib_device()) {
        float tab[8][128];
        paint_cos_tables(tab);
        PAINT_TRY(hipMemcpyToSymbol(HIP_SYMBOL(g_cos_tab), tab, sizeof tab, 0, hipMemcpyHostToDevice));
        for (auto &e : g_ev) {  // events of this device
            if (e) (void)hipEventDestroy(e);
            e = nullptr;
            PAINT_TRY(hipEventCreate(&e));
        }
        g_tab_device = lib_device();
    }
    const int bpp = fmt == MB3D_HIP_PAINT_U8 ? 3 : fmt == MB3D_HIP_PAINT_U16 ? 6 : 12;
    const uint32_t *d_g = gbuf;
    unsigned char *d_rgb = static_cast<unsigned char *>(rgb);
    if (!on_device) {
        const size_t gb = (size_t)npix * 20, ob = (size_t)npix * bpp;
        int status = 0;
        char *work = static_cast<char *>(lib_work(gb + ob + 256, &status));
        if (!work) return status;
        d_g = reinterpret_cast<const uint32_t *>(work);
        d_rgb = reinterpret_cast<unsigned char *>(work + ((gb + 255) & ~(size_t)255));
        PAINT_TRY(hipMemcpyAsync(const_cast<uint32_t *>(d_g), gbuf, gb, hipMemcpyHostToDevice, st));
    }
    PAINT_TRY(hipEventRecord(g_ev[0], st));
    const long long blocks = (npix + kPaintBlock - 1) / kPaintBlock;
    const int grid = (int)std::min<long long>(blocks, (long long)std::max(lib_cus(), 1) * 8);
    if (fmt == MB3D_HIP_PAINT_U8)
        hipLaunchKernelGGL(k_paint<MB3D_HIP_PAINT_U8>, dim3(grid), dim3(kPaintBlock), 0, st, *cam, *lv, d_g, x0, y0, w,
                           npix, d_rgb);
    else if (fmt == MB3D_HIP_PAINT_U16)
        hipLaunchKernelGGL(k_paint<MB3D_HIP_PAINT_U16>, dim3(grid), dim3(kPaintBlock), 0, st, *cam, *lv, d_g, x0, y0,
                           w, npix, d_rgb);
    else
        hipLaunchKernelGGL(k_paint<MB3D_HIP_PAINT_F32>, dim3(grid), dim3(kPaintBlock), 0, st, *cam, *lv, d_g, x0, y0,
                           w, npix, d_rgb);
    PAINT_TRY(hipGetLastError());
    PAINT_TRY(hipEventRecord(g_ev[1], st));
    if (!on_device) PAINT_TRY(hipMemcpyAsync(rgb, d_rgb, (size_t)npix * bpp, hipMemcpyDeviceToHost, st));
    PAINT_TRY(hipStreamSynchronize(st));
    float ms = 0.0f;
    PAINT_TRY(hipEventElapsedTime(&ms, g_ev[0], g_ev[1]));
    thread_counters().paint_ms = ms;
    return MB3D_HIP_OK;
}

}  // namespace

}  // namespace mb3d

using namespace mb3d;

extern "C" int32_t mb3d_hip_paint(const mb3d_hip_paint_camera *camera, const mb3d_hip_light_vals *light_vals,
                                  const uint32_t *gbuf, int32_t x0, int32_t y0, int32_t w, int32_t h, int32_t format,
                                  void *rgb, int32_t on_device) {
    try {
        std::lock_guard<std::mutex> lk(lib_mutex());
        return paint_locked(camera, light_vals, gbuf, x0, y0, w, h, format, rgb, on_device);
    } catch (...) {
        return set_error(MB3D_HIP_ERR_DEVICE, "unexpected failure in mb3d_hip_paint");
    }
}
