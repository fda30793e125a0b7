// paint_oracle.cpp -- TEST INFRASTRUCTURE, not product code.
//
// The CPU painter of the oracle: a restatement of mb3d-rs's lighting and
// painting of the G-buffer, used only by the tests as the checker of the
// HIP painter (and of the product's set-up, field by field).
//
//   orc_lighting_default   Lighting::default            src/lighting.rs:290-392
//   orc_paint_setup        paint_camera                 src/render.rs:2340-2380
//                          light_vals                   src/render.rs:2285-2319
//                          LightVals::build             src/lighting.rs:978-1252
//                          set_vol_light                src/lighting.rs:1273-1283
//                          rotate_object_lights         src/lighting.rs:1350-1381
//                          place_lights                 src/lighting.rs:1285-1320
//                          calc_z_pos_for_light         src/lighting.rs:834-897
//   orc_cos_tab_val        cos_tabs, cos_tab_val        src/lighting.rs:464-541
//   orc_paint              pixel_color_s                src/lighting.rs:1739-1760
//                          plv_for_pixel                src/lighting.rs:1762-1798
//                          shade                        src/lighting.rs:1800-2150
//                          calc_colors_alpha, palette_place, calc_colors_inside
//                                                       src/lighting.rs:1560-1705
//                          pos_light_shape              src/lighting.rs:899-941
//   orc_gbuf_pack/unpack   gbuffer::pack / unpack       src/gbuffer.rs:60-98
//
// f32 where the reference computes in f32 (SVec = [f32; 3]), f64 where it
// computes in f64; glibc's sinf/cosf/asinf as the reference's f32 methods
// call them.  Built with -ffp-contract=off: no fused operations.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

#include "mb3d_oracle.h"

namespace {

const double PI = 3.14159265358979323846;

struct SVec {
    float v[3];
    float &operator[](int i) { return v[i]; }
    float operator[](int i) const { return v[i]; }
};

SVec sv(float a, float b, float c) { return SVec{{a, b, c}}; }
SVec sv_add(SVec a, SVec b) { return sv(a[0] + b[0], a[1] + b[1], a[2] + b[2]); }
SVec sv_sub(SVec a, SVec b) { return sv(a[0] - b[0], a[1] - b[1], a[2] - b[2]); }
SVec sv_scale(SVec a, float s) { return sv(a[0] * s, a[1] * s, a[2] * s); }
SVec sv_mul(SVec a, SVec b) { return sv(a[0] * b[0], a[1] * b[1], a[2] * b[2]); }
void sv_add_w(SVec &a, SVec b, float w) {
    a[0] += b[0] * w;
    a[1] += b[1] * w;
    a[2] += b[2] * w;
}
float sv_dot(SVec a, SVec b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
// LinInterpolate2SVecs: sv1 * w1 + sv2 * (1 - w1)
SVec lerp(SVec sv1, SVec sv2, float w1) {
    return sv(sv2[0] + w1 * (sv1[0] - sv2[0]), sv2[1] + w1 * (sv1[1] - sv2[1]), sv2[2] + w1 * (sv1[2] - sv2[2]));
}
SVec col255(const uint8_t c[3]) { return sv((float)c[0], (float)c[1], (float)c[2]); }
SVec col1(const uint8_t c[3]) { return sv_scale(col255(c), 1.0f / 255.0f); }
float luma(SVec v) { return v[0] * 0.3f + v[1] * 0.59f + v[2] * 0.11f; }
SVec from_arr(const float a[3]) { return sv(a[0], a[1], a[2]); }
void to_arr(SVec s, float a[3]) {
    for (int k = 0; k < 3; k++) a[k] = s[k];
}
// SqrDistSV
float sqr_dist_sv(SVec a, SVec b) {
    SVec c = sv(a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]);
    return sv_dot(c, c);
}

float clampf_rs(float v, float lo, float hi) {
    if (v < lo) v = lo;
    if (v > hi) v = hi;
    return v;
}
double clampd_rs(double v, double lo, double hi) {
    if (v < lo) v = lo;
    if (v > hi) v = hi;
    return v;
}
int32_t as_i32(float v) {  // Rust's saturating cast
    if (std::isnan(v)) return 0;
    if (v >= 2147483648.0f) return INT32_MAX;
    if (v < -2147483648.0f) return INT32_MIN;
    return (int32_t)v;
}
// f32 powi as LLVM expands it: res starts at the first set bit's square
float powi32(float x, int n) {
    float r = 1.0f;
    bool first = true;
    float cur = x;
    for (unsigned u = (unsigned)n; u; u >>= 1) {
        if (u & 1u) {
            r = first ? cur : r * cur;
            first = false;
        }
        if (u >> 1) cur = cur * cur;
    }
    return r;
}
float fast_int_pow(float base, int expo) {  // src/math.rs:273-285
    if (base < 0.0f) return 0.0f;
    float r = base;
    int i = expo >> 1;
    while (i > 0) {
        r *= r;
        i >>= 1;
    }
    return r;
}
void make_spline_coeff(double xs, float w[4]) {  // src/math.rs:263-270
    const double r3 = xs * xs * xs / 6.0;
    const double r0 = 1.0 / 6.0 + 0.5 * xs * (xs - 1.0) - r3;
    const double r2 = xs + r0 - 2.0 * r3;
    const double r1 = 1.0 - r0 - r2 - r3;
    w[0] = (float)r0, w[1] = (float)r1, w[2] = (float)r2, w[3] = (float)r3;
}

// ---- cos_tabs (src/lighting.rs:464-499)
struct CosTabs {
    float diff[8][128];
    CosTabs() {
        for (int i = 0; i < 128; i++) {
            const double d = 1.0 - ((double)i - 2.0) / 60.0;
            float v0;
            if (d > 0.15)
                v0 = (float)((d - 0.08) * 1.0869565);
            else if (d <= 0.0)
                v0 = 0.0f;
            else
                v0 = (float)std::pow(d, std::fmax(1.0, (0.505 - d) * 3.8));
            diff[0][i] = v0;
            const double dc = std::fmax(d, 0.0);
            diff[1][i] = (float)(dc * dc);
            diff[2][i] = (float)(d * 0.5 + 0.5);
            diff[3][i] = (float)((d * 0.5 + 0.5) * (d * 0.5 + 0.5));
        }
        for (int k = 0; k < 4; k++) {
            std::vector<float> tmp(128);
            for (int j = 0; j < 128; j++) tmp[j] = std::sqrt(std::fmax(diff[k][j], 0.0f));
            for (int j = 0; j < 128; j++) {
                double e = 0.0;
                for (int i = 0; i <= 60; i++) {
                    const long l = std::labs((long)j + i - 30);
                    if (l < 128) e += (double)tmp[l];
                }
                const double a = e * 0.011 + (e * 0.007) * (e * 0.007);
                diff[k + 4][j] = (float)(a * a);
            }
        }
    }
};
const CosTabs &tabs() {
    static const CosTabs t;
    return t;
}

float cos_tab_val(int tnr, float dotp, float rough, bool sqr) {
    const CosTabs &tb = tabs();
    float t = 62.0f - 60.0f * dotp;
    int ip = as_i32(std::trunc(t)) - 1;
    if (ip < 0) {
        ip = 0;
        t = 0.0f;
    } else if (ip > 124) {
        ip = 124;
        t = 1.0f;
    } else {
        t = t - std::trunc(t);
    }
    float w[4];
    make_spline_coeff((double)t, w);
    float a = 0.0f, b = 0.0f;
    for (int k = 0; k < 4; k++) a += tb.diff[tnr][ip + k] * w[k];
    for (int k = 0; k < 4; k++) b += tb.diff[tnr + 4][ip + k] * w[k];
    if (sqr) {
        a *= a;
        b *= b;
    }
    return a + rough * (b - a);
}

// ---- PaintCamera (src/lighting.rs:726-815)
SVec view_vec(const mb3d_hip_paint_camera &c, float x_pos, float y_pos, float aspect) {
    const float cx = (c.x_off - x_pos) * c.fov * aspect;
    const float cy = (y_pos - 0.5f) * c.fov;
    switch (c.planar) {
        case 1: {
            SVec v = sv(-cx, cy, c.pl_optic_z);
            const float n = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
            return sv_scale(v, 1.0f / n);
        }
        case 2: return sv(-std::sin(cx) * std::cos(cy), std::sin(cy), std::cos(cy) * std::cos(cx));
        case 3: {
            const double xa = (double)cy, ya = (double)cx;
            const double t = std::sqrt(xa * xa + ya * ya);
            if (t < 1e-12) return sv(0.0f, 0.0f, 1.0f);
            const double tc = std::fmin(t, PI);
            const double s = std::sin(tc) / t;
            return sv((float)(-ya * s), (float)(xa * s), (float)std::cos(tc));
        }
        default: {
            SVec v = sv(-std::sin(cx), std::sin(cy), std::cos(cy) * std::cos(cx));
            const float n = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
            return sv_scale(v, 1.0f / n);
        }
    }
}
SVec to_abs(const mb3d_hip_paint_camera &c, SVec v) {
    return sv(v[0] * c.m[0][0] + v[1] * c.m[1][0] + v[2] * c.m[2][0], v[0] * c.m[0][1] + v[1] * c.m[1][1] + v[2] * c.m[2][1],
              v[0] * c.m[0][2] + v[1] * c.m[1][2] + v[2] * c.m[2][2]);
}
SVec to_view(const mb3d_hip_paint_camera &c, SVec v) {
    return sv(sv_dot(from_arr(c.m[0]), v), sv_dot(from_arr(c.m[1]), v), sv_dot(from_arr(c.m[2]), v));
}
SVec cam_pos(const mb3d_hip_paint_camera &c, float x, float y) {
    if (c.planar == 2) return from_arr(c.start_pos);
    return sv_add(sv_add(from_arr(c.start_pos), sv_scale(from_arr(c.y_add), y)), sv_scale(from_arr(c.x_add), x));
}

double signum(double x) { return std::isnan(x) ? x : (std::signbit(x) ? -1.0 : 1.0); }

float calc_z_pos_for_light(const mb3d_hip_paint_camera &cam, SVec lpos, bool positional, const double view_z[3]) {
    if (!positional) return 0.0f;
    const float w = (float)cam.width, h = (float)cam.height;
    const float aspect = w / h;
    auto plv = [&](float mx, float my, SVec &sp, SVec &ab) {
        sp = sv_add(sv_add(from_arr(cam.start_pos), sv_scale(from_arr(cam.y_add), my)), sv_scale(from_arr(cam.x_add), mx));
        ab = to_abs(cam, view_vec(cam, mx / w, my / h, aspect));
    };
    auto dist = [&](SVec ab, SVec sp) -> double {
        SVec v1 = sv_sub(lpos, sp);
        SVec v2 = sv(v1[1] * (v1[2] - ab[2]) - v1[2] * (v1[1] - ab[1]), v1[2] * (v1[0] - ab[0]) - v1[0] * (v1[2] - ab[2]),
                     v1[0] * (v1[1] - ab[1]) - v1[1] * (v1[0] - ab[0]));
        return (double)std::sqrt(sv_dot(v2, v2));
    };
    const double dmul = 0.001 / cam.step_width;
    float x = w * 0.5f, y = h * 0.5f;
    SVec sp, ab;
    plv(x, y, sp, ab);
    for (int it = 0; it < 21; it++) {
        plv(x, y, sp, ab);
        const double d = dist(ab, sp) * dmul;
        if (d < 1e-4) break;
        SVec s1, a1, s2, a2;
        plv(x + 0.01f, y, s1, a1);
        const double dx = d - dist(a1, s1) * dmul;
        plv(x, y + 0.01f, s2, a2);
        const double dy = d - dist(a2, s2) * dmul;
        const double dd = std::sqrt(dx * dx + dy * dy) + 1e-16;
        sp = s2;
        ab = a2;
        if (dd * 2.0 > d) break;
        x += (float)(signum(dx) * d * 4e-3 / dd);
        y += (float)(signum(dy) * d * 4e-3 / dd);
    }
    double z;
    if (std::fabs(ab[0]) > 0.5f)
        z = (double)((lpos[0] - sp[0]) / ab[0]);
    else if (std::fabs(ab[1]) > 0.5f)
        z = (double)((lpos[1] - sp[1]) / ab[1]);
    else
        z = (double)((lpos[2] - sp[2]) / ab[2]);
    const double dz = cam.zz_stmit_dif;
    double dv[3];
    for (int k = 0; k < 3; k++) dv[k] = (double)lpos[k] - view_z[k] * dz;
    const double d = dv[0] * view_z[0] + dv[1] * view_z[1] + dv[2] * view_z[2];
    if (std::fabs(d) > std::fabs(z)) z = d;
    return (float)std::fmax(z, 0.0);
}

void pos_light_shape(float &flux, float &transp, int func, bool pos, bool sqr) {
    const float tmp_r = transp;
    const float rtmp_r = 1.0f / tmp_r;
    switch (func) {
        case 4: {
            const float t = flux * rtmp_r;
            transp = (t * t) * (t * t);
            flux = 7e-3f * (tmp_r - flux) * rtmp_r * rtmp_r;
            break;
        }
        case 2: {
            transp = 1.0f;
            const float f = 1.0f - std::sqrt(flux * rtmp_r);
            flux = 5e-2f * powi32(std::fmax(f, 0.0f), 50) * rtmp_r;
            if (pos) flux *= 1.5f;
            break;
        }
        case 6: {
            const float f = flux * 1.05f * rtmp_r;
            if (f < 1.0f) {
                transp = pos ? std::fmax(f - 0.95f, 0.0f) * 10.0f : 0.7f;
                flux = 4e-3f * std::sqrt(1.05f - f * f) * rtmp_r;
                if (!pos) flux *= 0.7f;
            } else if (pos) {
                transp = (f - 0.95f) * 10.0f;
                flux = 17.8e-3f * (1.05f - f) * rtmp_r;
            } else {
                transp = (f - 1.0f) * 6.0f + 0.7f;
                flux = 12.5e-3f * (1.05f - f) * rtmp_r;
            }
            break;
        }
        default: {
            transp = 1.0f;
            const float t = flux * rtmp_r;
            const float k = sqr ? 40.0f : 25.0f;
            flux = (1.0f - t) * 0.1f / powi32((t + 0.0004f) * k, 2) * rtmp_r;
        }
    }
}

uint32_t convert_vlight(uint16_t w) {  // src/lighting.rs:944-948
    const uint32_t x = (uint32_t)(w & 0x3FF);
    return (x & 0x7F) << ((x >> 7) & 7);
}

// ---- the G-buffer record (src/gbuffer.rs:4-59 from the card words, src/gpu.rs:3452-3463)
struct SiLight {
    int16_t normal[3];
    uint32_t zpos_fine;
    uint16_t shadow, amb_shadow, si_gradient, otrap;
    uint32_t zpos() const { return zpos_fine >> 16; }
    bool is_ground() const { return otrap == 0xFFFF; }
};
SiLight from_words(const uint32_t *u) {
    SiLight s;
    s.normal[0] = (int16_t)(uint16_t)u[0];
    s.normal[1] = (int16_t)(uint16_t)(u[0] >> 16);
    s.normal[2] = (int16_t)(uint16_t)u[1];
    s.shadow = (uint16_t)(u[1] >> 16);
    s.zpos_fine = u[2];
    s.si_gradient = (uint16_t)u[3];
    s.otrap = (uint16_t)(u[3] >> 16);
    s.amb_shadow = (uint16_t)u[4];
    return s;
}

// ---- LightVals (src/lighting.rs:605-698) read from the shared layout
struct Lv {
    const mb3d_hip_light_vals &l;
    SVec amb_abs(SVec v) const {
        if (!l.has_amb_frame) return v;
        return sv(sv_dot(from_arr(l.amb_frame[0]), v), sv_dot(from_arr(l.amb_frame[1]), v), sv_dot(from_arr(l.amb_frame[2]), v));
    }
    // palette_place: (a, has_b, b, t)
    void palette_place(const SiLight &si, float idif0, int &a, bool &two, int &b, float &t) const {
        const float irf = l.col_on_otrap ? (float)(si.otrap & 0x7FFF) : (float)si.si_gradient;
        int32_t ir = (int32_t)std::round(clampf_rs(((irf - l.s_c_start) * l.s_c_mul + idif0) * 16384.0f, -1e9f, 1e9f));
        two = false;
        if (l.col_cycling) {
            ir &= 32767;
        } else if (ir < 0) {
            a = 0;
            return;
        } else if (ir >= l.col_pos[9]) {
            a = 9;
            return;
        }
        int il2 = 5;
        if (l.col_pos[il2] < ir) {
            for (;;) {
                il2 += 1;
                if (il2 == 10 || l.col_pos[il2] >= ir) break;
            }
        } else {
            while (il2 > 1 && l.col_pos[il2 - 1] >= ir) il2 -= 1;
        }
        if (l.no_col_ipol) {
            a = il2 - 1;
            return;
        }
        const int il1 = il2 - 1;
        a = il1;
        two = true;
        b = il2 > 9 ? 0 : il2;
        t = (float)(ir - l.col_pos[il1]) * l.s_c_div[il1];
    }
    void calc_colors(const SiLight &si, float idif0, SVec &dif, SVec &spe) const {
        if (si.is_ground() && l.has_ground) {
            dif = from_arr(l.ground_dif);
            spe = from_arr(l.ground_spe);
            return;
        }
        int a = 0, b = 0;
        bool two = false;
        float t = 0.0f;
        palette_place(si, idif0, a, two, b, t);
        if (!two) {
            dif = from_arr(l.col_dif[a]);
            spe = from_arr(l.col_spe[a]);
            return;
        }
        dif = lerp(from_arr(l.col_dif[b]), from_arr(l.col_dif[a]), t);
        spe = lerp(from_arr(l.col_spe[b]), from_arr(l.col_spe[a]), t);
    }
    void calc_colors_inside(const SiLight &si, float idif0, SVec &dif, SVec &spe3) const {
        int32_t ir = as_i32(std::round((((float)si.si_gradient - l.s_ci_start) * l.s_ci_mul + idif0) * 16384.0f));
        float spe;
        if (l.col_cycling) ir &= 32767;
        if (!l.col_cycling && ir < 0) {
            dif = from_arr(l.col_int[0]);
            spe = l.col_int_spec[0];
        } else if (!l.col_cycling && ir > l.icol_pos[3]) {
            dif = from_arr(l.col_int[3]);
            spe = l.col_int_spec[3];
        } else {
            int il2 = 1;
            while (il2 < 4 && l.icol_pos[il2] < ir) il2 += 1;
            if (l.no_col_ipol) {
                dif = from_arr(l.col_int[il2 - 1]);
                spe = l.col_int_spec[il2 - 1];
            } else {
                const int il1 = il2 - 1;
                const int j = il2 & 3;
                const float t = (float)(ir - l.icol_pos[il1]) * l.s_ic_div[il1];
                dif = lerp(from_arr(l.col_int[j]), from_arr(l.col_int[il1]), t);
                spe = l.col_int_spec[j] + t * (l.col_int_spec[il1] - l.col_int_spec[j]);
            }
        }
        spe3 = sv(spe, spe, spe);
    }
};

// plv_for_pixel + shade (own pixel, default ShadeOpts)
SVec shade(const Lv &lvw, const SiLight &si, int x, int y, const mb3d_hip_paint_camera &cam) {
    const mb3d_hip_light_vals &L = lvw.l;
    // plv_for_pixel
    const float x_pos = (float)(x + 1) / (float)cam.width;
    const float y_pos = (float)y / (float)cam.height;
    const SVec view = view_vec(cam, x_pos, y_pos, cam.aspect);
    float ys;
    if (L.depth_func == 1)
        ys = y_pos * y_pos;
    else if (L.depth_func == 0)
        ys = y_pos;
    else
        ys = std::sqrt(std::fmax(y_pos, 0.0f));
    const uint32_t zpos_word = si.zpos();
    double z1;
    if (zpos_word > 32767) {
        const double q = 8388352.0 / cam.zc_mul + 1.0;
        z1 = (q * q - 1.0) * cam.step_width / cam.zcorr;
    } else {
        const double q = (double)(8388352LL - (int64_t)(si.zpos_fine >> 8)) / cam.zc_mul + 1.0;
        z1 = (q * q - 1.0) * cam.step_width / cam.zcorr;
    }
    const float z_pos = (float)(z1 + cam.zz_stmit_dif);
    const SVec abs_view = to_abs(cam, view);
    const SVec cpos = cam_pos(cam, (float)x, (float)y);
    const SVec obj_pos = sv_add(cpos, sv_scale(abs_view, (float)z1));
    SVec dep_c = lerp(from_arr(L.depth_col2), from_arr(L.depth_col), ys);

    // shade
    if (L.amb_rel_obj) {
        const float a = std::asin(clampf_rs(lvw.amb_abs(abs_view)[1], -1.0f, 1.0f));
        float w;
        if (L.amb_fov > 0.0f) {
            const float yy = clampf_rs(a / L.amb_fov + 0.5f, 0.0f, 1.0f);
            w = L.depth_func == 1 ? yy * yy : (L.depth_func == 0 ? yy : std::sqrt(yy));
        } else {
            w = a * (float)(1.0 / PI) + 0.5f;
        }
        dep_c = lerp(from_arr(L.depth_col2), from_arr(L.depth_col), w);
    }
    const bool sqr = L.sqr != 0;
    SVec out;
    float dtmp;
    if (zpos_word < 32768) {
        const float d_rough = (float)(si.zpos_fine & 0xFF) * L.s_roughness_factor;
        const float shadow = si.amb_shadow >= 16383 ? 1.0f : (float)si.amb_shadow / 16383.0f;
        float d_amb_sh;
        if (L.s_amb_shad < 1.0f) {
            d_amb_sh = 1.0f - L.s_amb_shad * shadow;
        } else {
            const float a = 1.0f - shadow;
            d_amb_sh = a + (L.s_amb_shad - 1.0f) * (a * a - a);
        }
        if (sqr) d_amb_sh = (d_amb_sh * d_amb_sh + d_amb_sh) * 0.5f;
        const SVec lns = sv((float)si.normal[0] * 3.0518509476e-5f, (float)si.normal[1] * 3.0518509476e-5f,
                            (float)si.normal[2] * 3.0518509476e-5f);
        SVec li_dif = sv(0, 0, 0), li_spe = sv(0, 0, 0), li_amb = sv(0, 0, 0);
        const float d_fog = L.s_diffuse_shadowing * (d_amb_sh - 1.0f) + 1.0f;
        for (int i = 0; i < L.n_lights; i++) {
            const mb3d_hip_light_val &lv = L.lights[i];
            const bool soft = lv.hs_mask == 0xFFFF;
            const bool no_hs = soft || (si.shadow & (uint16_t)lv.hs_mask) == 0 || !lv.hs_calced;
            const float soft_mul = soft ? (float)(si.shadow >> 10) * (1.0f / 63.0f) : 1.0f;
            SVec ln;
            float att;
            if (lv.positional) {
                const SVec v = sv_sub(from_arr(lv.ln), obj_pos);
                const float d2 = sv_dot(v, v);
                if (d2 > lv.lmax_l) continue;
                ln = to_view(cam, sv_scale(v, 1.0f / std::sqrt(d2)));
                att = 1.0f / (d2 + 1e-30f);
            } else {
                ln = from_arr(lv.ln);
                att = 1.0f;
            }
            const float dot = lns[0] * ln[0] + lns[1] * ln[1] + lns[2] * ln[2];
            float d = cos_tab_val(lv.diff_func, dot, d_rough, sqr) * att;
            sv_add_w(li_amb, from_arr(lv.col), d);
            if (no_hs) {
                d *= lv.sub_amb_sh ? d_amb_sh : d_fog;
                d *= soft_mul;
                sv_add_w(li_dif, from_arr(lv.col), d);
            }
            const float d2 = 2.0f * (lns[0] * view[0] + lns[1] * view[1] + lns[2] * view[2]);
            const float ds = ln[0] * (view[0] - lns[0] * d2) + ln[1] * (view[1] - lns[1] * d2) + ln[2] * (view[2] - lns[2] * d2);
            if (ds > 0.0f) {
                const float pf = (float)lv.pow_func;
                float t2 = (att + std::fmin(d_rough * 2.0f, 1.0f) * (1.0f / pf - att)) * L.s_spec;
                if (t2 > 0.0f) {
                    sv_add_w(li_amb, from_arr(lv.col), t2 / pf);
                    if (no_hs) {
                        t2 *= fast_int_pow(ds, lv.pow_func);
                        t2 *= lv.sub_amb_sh ? d_amb_sh : d_fog;
                        t2 *= soft_mul;
                        sv_add_w(li_spe, from_arr(lv.col), t2);
                    }
                }
            }
        }
        const float idif0 = L.s_col_zmul * z_pos;
        SVec idif, ispe;
        if (si.si_gradient > 32767 && !si.is_ground())
            lvw.calc_colors_inside(si, idif0, idif, ispe);
        else
            lvw.calc_colors(si, idif0, idif, ispe);
        const float ny = L.amb_rel_obj ? lvw.amb_abs(to_abs(cam, lns))[1] : lns[1];
        const float t2 = (ny * 0.5f + 0.5f) * d_amb_sh;
        const float t1 = d_amb_sh - t2;
        SVec iamb = sv_add(sv_scale(from_arr(L.amb_col), t1), sv_scale(from_arr(L.amb_col2), t2));
        iamb = sv_mul(iamb, idif);
        const SVec dif_s = sv_scale(idif, L.s_diff);
        const SVec total = sv_add(sv_add(iamb, sv_mul(dif_s, li_dif)), sv_mul(ispe, li_spe));
        if (!L.ex_mode) {
            const SVec l3 = sv_add(iamb, sv_scale(sv_mul(dif_s, li_amb), d_amb_sh * (sqr ? 0.11f : 0.3f)));
            const float st = std::fmax(1.0f - d_amb_sh, 0.0f);
            const float st2 = st * st * st * st;
            SVec refl;
            for (int k = 0; k < 3; k++) refl[k] = idif[k] * st2 * L.s_ind_light_reflect + (st - st2);
            out = sv_add(total, sv_mul(l3, sv_mul(idif, refl)));
        } else {
            float st = std::fmax(1.0f - d_amb_sh, 0.0f);
            st = st * st * L.s_ind_light_reflect;
            const float st2 = luma(total) + 4.0f;
            out = sv_add(total, sv_scale(sv_mul(total, idif), st));
            out = sv_scale(out, (st2 + (st2 + 128.0f) * (st + 0.11f) * 0.04f) / (std::fmax(luma(out), 0.0f) + 3.0f));
        }
        dtmp = std::fmax(((float)zpos_word - 28000.0f) * L.s_depth + 1.0f, 0.0f);
    } else {
        out = dep_c;
        dtmp = std::fmax(1.0f - (float)(60768 - (int32_t)zpos_word) * L.s_depth, 0.0f);
    }
    if (sqr) {
        if (dtmp < 1.0f) {
            float t = (1.0f - dtmp) * (1.0f - dtmp);
            if (L.far_fog) t *= t;
            dtmp = 1.0f - t;
        }
    } else if (L.far_fog && dtmp < 1.0f) {
        dtmp = 1.0f - (1.0f - dtmp) * (1.0f - dtmp);
    }
    out = sv_scale(out, dtmp);
    const float ir = L.vol_light ? (float)convert_vlight(si.shadow) : (float)(si.shadow & 0x3FF);
    float dfog = (ir - L.s_shad - L.s_shad_zmul * z_pos) * L.s_shad_gr;
    if (L.dfog_options & 2) dfog = std::fmax(dfog, 0.0f);
    float dtmp3 = std::fmin(ir * L.s_dyn_fog_mul, 1.0f) * dfog;
    out = sv_add(out, sv_scale(dep_c, std::fmax(1.0f - dtmp, 0.0f)));
    if (L.dfog_options & 1) {
        dfog = clampf_rs(dfog, 0.0f, 1.0f);
        dtmp3 = clampf_rs(dtmp3, 0.0f, 1.0f);
        out = sv_scale(out, 1.0f - dfog);
    }
    out = sv_add(out, sv_add(sv_scale(from_arr(L.dyn_fog_col), dfog - dtmp3), sv_scale(from_arr(L.dyn_fog_col2), dtmp3)));
    // visible lights
    const int32_t il1 = zpos_word < 32768 ? 32768 - (int32_t)zpos_word : (int32_t)zpos_word;
    const float d_rough_v = 1.0f / (float)il1;
    for (int i = 0; i < L.n_lights; i++) {
        const mb3d_hip_light_val &lv = L.lights[i];
        if (lv.visible == 0) continue;
        const bool pos = lv.positional != 0;
        bool behind = false;
        float flux;
        if (pos) {
            flux = sqr_dist_sv(sv_sub(cpos, from_arr(lv.ln)), abs_view);
        } else {
            if (zpos_word < 32768) continue;
            flux = std::fmax(1.0f - sv_dot(from_arr(lv.ln), view), 0.0f);
        }
        float transp = lv.lmax_l * 1e-8f;
        if (flux >= transp) continue;
        if (pos) {
            if (sv_dot(sv_sub(from_arr(lv.ln), cpos), abs_view) < 0.0f) continue;
            if (zpos_word < 32768) {
                const float t = -std::sqrt(transp - flux);
                if (std::fabs(abs_view[0]) > 0.5f)
                    behind = t > (obj_pos[0] - lv.ln[0]) / abs_view[0];
                else if (std::fabs(abs_view[1]) > 0.5f)
                    behind = t > (obj_pos[1] - lv.ln[1]) / abs_view[1];
                else
                    behind = t > (obj_pos[2] - lv.ln[2]) / abs_view[2];
            }
        }
        if (behind) continue;
        pos_light_shape(flux, transp, lv.visible, pos, sqr);
        float fog_at = 1.0f - std::fmax(1.0f + (lv.pos_lp - 28000.0f) * L.s_depth, 0.0f);
        if (sqr && fog_at > 0.0f) fog_at *= fog_at;
        if (L.far_fog && fog_at > 0.0f) fog_at *= fog_at;
        flux *= 1.0f - fog_at * 0.9f;
        if (lv.visible == 2 || lv.visible == 8) {
            out = sv_add(out, sv_scale(from_arr(lv.col), flux));
        } else {
            fog_at = std::fmax(fog_at, 0.0f);
            const float il = L.vol_light ? (float)convert_vlight(si.shadow) : (float)(si.shadow & 0x3FF);
            const float k = il * (32768.0f - lv.pos_lp) * d_rough_v;
            const float df = (k - L.s_shad - L.s_shad_zmul * (lv.pos_z + (float)cam.zz_stmit_dif)) * L.s_shad_gr *
                             (1.0f - transp) * (1.0f - fog_at);
            const float damb = std::fmin(k * L.s_dyn_fog_mul, 1.0f) * df;
            out = sv_add(sv_add(sv_scale(out, transp), sv_scale(dep_c, fog_at * (1.0f - transp))),
                         sv_add(sv_add(sv_scale(from_arr(L.dyn_fog_col), df - damb), sv_scale(from_arr(L.dyn_fog_col2), damb)),
                                sv_scale(from_arr(lv.col), flux)));
        }
    }
    return out;
}

void pixel_color(const Lv &lvw, const SiLight &si, int x, int y, const mb3d_hip_paint_camera &cam, int fmt, uint8_t *dst) {
    const mb3d_hip_light_vals &L = lvw.l;
    const SVec o = shade(lvw, si, x, y, cam);
    const bool hdr = fmt == MB3D_HIP_PAINT_F32;
    const float top = hdr ? 3.40282347e38f : 255.0f;
    SVec c;
    for (int k = 0; k < 3; k++)
        c[k] = L.sqr ? std::sqrt(clampf_rs(o[k] * 255.0f, 0.0f, top * top)) : clampf_rs(o[k], 0.0f, top);
    if (L.gamma_h != 0) {
        SVec g;
        for (int k = 0; k < 3; k++) g[k] = L.gamma_h > 0 ? std::sqrt(c[k] * 255.0f) : c[k] * c[k] / 255.0f;
        c = lerp(g, c, L.s_gamma);
    }
    for (int k = 0; k < 3; k++) {
        if (fmt == MB3D_HIP_PAINT_U8) {
            const float v = std::round(c[k]);
            dst[k] = std::isnan(v) || v <= 0.0f ? 0 : (v >= 255.0f ? 255 : (uint8_t)v);
        } else if (fmt == MB3D_HIP_PAINT_U16) {
            const float v = std::round(c[k] * 257.0f);
            const uint16_t u = std::isnan(v) || v <= 0.0f ? 0 : (v >= 65535.0f ? 65535 : (uint16_t)v);
            std::memcpy(dst + 2 * k, &u, 2);
        } else {
            const float v = c[k] / 255.0f;
            std::memcpy(dst + 4 * k, &v, 4);
        }
    }
}

void tcol(uint32_t c, uint8_t o[3]) {  // Delphi TColor $BBGGRR
    o[0] = (uint8_t)(c & 0xFF);
    o[1] = (uint8_t)((c >> 8) & 0xFF);
    o[2] = (uint8_t)((c >> 16) & 0xFF);
}

}  // namespace

extern "C" {

void orc_lighting_default(mb3d_hip_lighting *l) {
    std::memset(l, 0, sizeof *l);
    const uint32_t cols[9] = {5873889, 8837614, 0x8B491D, 2988346, 12248958, 0xFFC49F, 11287584, 14579248, 7481121};
    for (int i = 0; i < 3; i++) {
        l->palette[i].position = (uint16_t)(i * 10922);
        tcol(cols[i * 3], l->palette[i].diffuse);
        tcol(cols[i * 3 + 1], l->palette[i].specular);
    }
    for (int i = 3; i < 10; i++) {
        l->palette[i] = l->palette[0];
        l->palette[i].position = (uint16_t)(32200 + (i - 3) * 84);
    }
    for (int i = 0; i < 3; i++) {
        l->interior_pos[i] = (uint16_t)(i * 10922);
        tcol(cols[i * 3], l->interior_col[i]);
    }
    l->interior_pos[3] = 32700;
    for (int k = 0; k < 3; k++) l->interior_col[3][k] = l->interior_col[0][k];
    for (int i = 0; i < MB3D_HIP_MAX_LIGHTS; i++) {  // defaultLight8
        mb3d_hip_light &li = l->lights[i];
        li.color[0] = li.color[1] = li.color[2] = 255;
        li.amplitude = 1.0f;
        li.spec_func = 4;
        li.hs_enabled = 1;
    }
    l->lights[0].on = 1;
    tcol(0xA0E8FF, l->lights[0].color);
    l->lights[0].x_angle = 1500.0 * PI / 16384.0;
    l->lights[0].y_angle = 5200.0 * PI / 16384.0;
    l->lights[0].spec_func = 53 & 7;
    l->lights[0].diff_func = (53 >> 4) & 3;
    tcol(2307911, l->lights[1].color);
    l->lights[1].x_angle = -2822.0 * PI / 16384.0;
    l->lights[1].y_angle = -7737.0 * PI / 16384.0;
    l->lights[1].spec_func = 16 & 7;
    l->lights[1].diff_func = (16 >> 4) & 3;
    tcol(0x8B491D, l->amb_top);
    tcol(0xFFC49F, l->amb_bottom);
    tcol(0x8B491D, l->depth_col);
    tcol(0xFFC49F, l->depth_col2);
    for (int k = 0; k < 3; k++) l->dyn_fog_col[k] = l->dyn_fog_col2[k] = 255;
    l->fog_offset = 128.0f;
    l->fog_far_offset = 128.0f;
    l->depth_fog = 27.0f;
    l->diffuse = 120.0f;
    l->dyn_fog = 53.0f;
    l->specular = 50.0f;
    l->ambient = 90.0f;
    l->color_start = 68.0f;
    l->color_end = 72.0f;
    l->amb_shadow = 53.0f;
    l->ind_light = 53.0f;
    l->gamma = 32.0f;
    l->interior_start = 60.0f;
    l->interior_end = 60.0f;
    for (int k = 0; k < 4; k++) l->interior_spec[k] = 160;
    l->fine_col_adj[0] = l->fine_col_adj[1] = 30;
    l->vol_light = -1;
}

int32_t orc_paint_setup(const mb3d_hip_scene *sc, const mb3d_hip_params *p, const mb3d_hip_lighting *lt,
                        mb3d_hip_paint_camera *cam_out, mb3d_hip_light_vals *lv_out) {
    // paint_camera (common optic, no stereo)
    mb3d_hip_paint_camera cam;
    std::memset(&cam, 0, sizeof cam);
    const int w = sc->width, h = sc->height;
    cam.width = w;
    cam.height = h;
    cam.fov = (float)(sc->fov_y * (PI / 180.0));
    const double d = clampd_rs((double)cam.fov * 0.5, 0.01, 1.5);
    double m[3][3];
    for (int i = 0; i < 3; i++) {  // normalise_matrix_to(1.0, &p.vgrads)
        const double s = 1.0 / std::sqrt(p->vgrads[i][0] * p->vgrads[i][0] + p->vgrads[i][1] * p->vgrads[i][1] +
                                         p->vgrads[i][2] * p->vgrads[i][2]);
        for (int j = 0; j < 3; j++) m[i][j] = p->vgrads[i][j] * s;
    }
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) cam.m[i][j] = (float)m[i][j];
    const double swp = p->step_width;
    const double zsd = sc->z_start - sc->mid[2];  // CalcParams::zz_stmit_dif, src/calc.rs:858
    const double spv[3] = {-0.5 * (double)w * swp, -0.5 * (double)h * swp, zsd};
    for (int k = 0; k < 3; k++) {
        cam.start_pos[k] = (float)(spv[0] * m[0][k] + spv[1] * m[1][k] + spv[2] * m[2][k]);
        cam.x_add[k] = (float)(m[0][k] * swp);
        cam.y_add[k] = (float)(m[1][k] * swp);
    }
    cam.aspect = (float)w / (float)h;
    cam.x_off = 0.5f;  // Scene::stereo_xoff without stereo
    cam.planar = 0;
    cam.pl_optic_z = (float)(std::cos(d) * d / std::sin(d));
    cam.zcorr = p->zcorr;
    cam.zc_mul = p->zc_mul;
    cam.step_width = p->step_width;
    cam.zz_stmit_dif = zsd;

    // LightVals::new (image_scale 1)
    mb3d_hip_light_vals L;
    std::memset(&L, 0, sizeof L);
    const bool sqr = lt->internal_gamma2 != 0;
    const float light_scale = sqr ? 1.5f : 1.0f;
    auto lcol = [&](const uint8_t c[3]) {
        return sqr ? sv((float)c[0] * (float)c[0] / 255.0f, (float)c[1] * (float)c[1] / 255.0f, (float)c[2] * (float)c[2] / 255.0f)
                   : col255(c);
    };
    auto pcol = [&](const uint8_t c[3]) {
        return sqr ? sv((float)c[0] * (float)c[0] * 0.0000153787f, (float)c[1] * (float)c[1] * 0.0000153787f,
                        (float)c[2] * (float)c[2] * 0.0000153787f)
                   : col1(c);
    };
    auto hs_calced = [&](int idx) { return ((lt->hs_calced >> idx) & 1) != 0; };
    const double zsdiv = sc->z_step_div;
    const double z_range = sc->z_end - sc->z_start;
    const double sw = 2.1345 / (sc->zoom * (double)sc->width);  // Scene::step_width (src/scene.rs:717-719)
    int n = 0;
    for (int idx = 0; idx < MB3D_HIP_MAX_LIGHTS; idx++) {
        const mb3d_hip_light &li = lt->lights[idx];
        if (!li.on) continue;
        const double dt = -li.x_angle, dt2 = li.y_angle;
        const double v[3] = {-std::sin(dt), std::sin(dt2), std::cos(dt2) * std::cos(dt)};
        const double nv = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        mb3d_hip_light_val &o = L.lights[n++];
        for (int k = 0; k < 3; k++) o.ln[k] = (float)(-v[k] / nv);
        const SVec raw = sv_scale(lcol(li.color), light_scale);
        SVec col;
        float lmax;
        if (li.positional) {
            col = sv_scale(raw, li.amplitude * 1.3f * light_scale);
            lmax = 800.0f * ((col[0] + col[1]) / 3.0f + 128.0f * li.amplitude);
        } else {
            col = sv_scale(raw, li.amplitude);
            lmax = 800.0f * (raw[0] + raw[1] + raw[2] + 128.0f) * li.amplitude;
        }
        if (li.visible == 8) lmax *= 5.0f;
        o.on = 1;
        to_arr(col, o.col);
        o.positional = li.positional ? 1 : 0;
        o.visible = li.visible;
        o.lmax_l = lmax;
        o.pow_func = 2 << (li.spec_func & 7);
        o.diff_func = (hs_calced(idx) && lt->hs_set_cos) ? 0 : (li.diff_func & 3);
        o.idx = idx;
        o.sub_amb_sh = hs_calced(idx) != (li.hs_enabled != 0) ? 1 : 0;
        o.hs_calced = hs_calced(idx) ? 1 : 0;
        o.hs_mask = (lt->hs_soft && hs_calced(idx)) ? 0xFFFF : (0x400 << idx);
    }
    L.n_lights = n;
    L.s_depth = lt->depth_fog * 0.8e-6f;
    float gr = (lt->dyn_fog - 53.0f) * (float)1.0 * (float)zsdiv * 0.00065f;
    float dtmp = 2.2f / (float)zsdiv;
    float mul = (float)zsdiv * 0.015f;
    float stmp = 128.0f;
    if (sc->dfog_on_it > 0) {
        dtmp *= 0.25f;
        gr *= 4.0f;
        mul *= 4.0f;
    } else {
        stmp = 137.0f;
    }
    L.s_shad_gr = gr;
    L.s_dyn_fog_mul = mul;
    L.s_shad = (stmp - std::sqrt(std::fmax(lt->fog_offset, 0.0f)) * 11.313708f) * dtmp * 0.28f;
    L.s_shad_zmul = dtmp * 0.7f / (float)z_range * (128.0f - std::sqrt(std::fmax(lt->fog_far_offset, 0.0f)) * 11.313708f);
    L.s_amb_shad = lt->amb_shadow / 53.0f;
    float cs = powi32((lt->color_start + 30.0f) / 90.0f, 2) * 32767.0f - 10900.0f;
    float cm = powi32((lt->color_end + 30.0f) / 90.0f, 2) * 32767.0f - 10900.0f - cs;
    if (lt->fine_col_adj_on) {
        const float a1 = lt->fine_col_adj[0], a2 = lt->fine_col_adj[1];
        const float dd = cs + cm * (a2 - 30.0f) * 0.0166666666666666f;
        cs += cm * (a1 - 30.0f) * 0.0166666666666666f;
        cm = dd - cs;
    }
    cm = std::fabs(cm) > 0.001f ? 2.0f / cm : (cm < 0.0f ? -2000.0f : 2000.0f);
    L.s_c_start = cs;
    L.s_c_mul = cm;
    const float dti = powi32(lt->interior_start, 4);
    L.s_ci_start = dti * 0.000158025f + 32768.0f;
    float cim = (powi32(lt->interior_end, 4) - dti) * 0.5f * 0.00031605f;
    if (std::fabs(cim) > 1e-10f) cim = 1.0f / cim;
    L.s_ci_mul = cim;
    const int32_t gi = (int32_t)std::round(lt->gamma);
    if (gi == 32) {
        L.gamma_h = 0;
        L.s_gamma = 0.0f;
    } else if (gi < 32) {
        L.gamma_h = -1;
        L.s_gamma = 1.0f - (float)gi / 32.0f;
    } else {
        L.gamma_h = 1;
        L.s_gamma = (float)(gi - 32) / 31.0f;
    }
    float amb_mul = lt->ambient / 90.0f;
    if (sqr) amb_mul = (amb_mul * amb_mul + amb_mul) * 0.5f * light_scale;
    for (int i = 0; i < 10; i++) {
        to_arr(pcol(lt->palette[i].diffuse), L.col_dif[i]);
        to_arr(pcol(lt->palette[i].specular), L.col_spe[i]);
        const uint8_t *s = lt->palette[i].specular;
        const uint8_t ab = lt->palette_alpha_set ? lt->palette_alpha[i] : std::max(s[0], std::max(s[1], s[2]));
        const float a = (float)ab;
        L.col_spe_a[i] = sqr ? a * a * 0.0000153787f : a / 255.0f;
        L.col_pos[i] = lt->palette[i].position;
    }
    for (int i = 0; i < 10; i++) {
        const int32_t j = i < 9 ? L.col_pos[i + 1] - L.col_pos[i] : 32767 - L.col_pos[9];
        L.s_c_div[i] = j > 1 ? 1.0f / (float)j : 1.0f;
    }
    for (int i = 0; i < 4; i++) {
        to_arr(pcol(lt->interior_col[i]), L.col_int[i]);
        L.icol_pos[i] = lt->interior_pos[i];
    }
    for (int i = 0; i < 4; i++) {
        const int32_t j = i < 3 ? L.icol_pos[i + 1] - L.icol_pos[i] : 32767 - L.icol_pos[3];
        L.s_ic_div[i] = j > 1 ? 1.0f / (float)j : 1.0f;
    }
    const float ind = std::round(clampf_rs(lt->ind_light - 53.0f, -128.0f, 127.0f));
    float sil = powi32((ind + 53.0f) * 0.022f, 2);
    float sd = lt->diffuse * 0.02f;
    float ss = std::fmax(lt->specular * 0.02f, 0.004f);
    if (sqr) {
        sil *= 0.5f;
        sd = (sd * sd + sd) * 0.5f;
        ss = (ss * ss + ss) * 0.5f;
    }
    L.s_ind_light_reflect = sil;
    L.s_diff = sd;
    L.s_spec = ss;
    for (int i = 0; i < 4; i++) {
        const float a = (float)lt->interior_spec[i];
        L.col_int_spec[i] = sqr ? a * a * 0.0000153787f : a / 255.0f;
    }
    to_arr(sv_scale(lcol(lt->amb_top), amb_mul), L.amb_col);
    to_arr(sv_scale(lcol(lt->amb_bottom), amb_mul), L.amb_col2);
    to_arr(lcol(lt->depth_col), L.depth_col);
    to_arr(lcol(lt->depth_col2), L.depth_col2);
    to_arr(sqr ? lcol(lt->dyn_fog_col) : col255(lt->dyn_fog_col), L.dyn_fog_col);
    to_arr(lcol(lt->dyn_fog_col2), L.dyn_fog_col2);
    L.s_col_zmul = (float)((double)lt->var_col_z * -0.005 / (sw * (double)sc->width));
    L.s_roughness_factor = lt->roughness / (255.0f * 255.0f);
    L.col_cycling = lt->color_cycling != 0;
    L.col_on_otrap = lt->color_on_otrap != 0;
    L.far_fog = lt->far_fog != 0;
    L.depth_func = lt->depth_func;
    L.s_diffuse_shadowing = lt->diffuse_shadowing;
    L.sqr = sqr;
    L.no_col_ipol = lt->no_col_ipol != 0;
    L.amb_rel_obj = lt->amb_rel_obj != 0;
    L.has_amb_frame = lt->has_amb_frame != 0;
    if (L.has_amb_frame) {
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) L.amb_frame[i][j] = (float)lt->amb_frame[i][j];
        L.amb_fov = (float)lt->amb_fov;
    }
    L.dfog_options = (lt->dfog_options & 3) == 3 ? 1 : (lt->dfog_options & 3);
    L.ex_mode = lt->ex_mode != 0;
    // light_vals: set_vol_light when the volumetric light's light is on
    if (lt->vol_light >= 0 && lt->vol_light < MB3D_HIP_MAX_LIGHTS && lt->lights[lt->vol_light].on) {
        L.vol_light = 1;
        L.s_dyn_fog_mul = 0.0005f;
        const float d50 = 50.0f;
        L.s_shad_gr = (lt->dyn_fog - 53.0f) * 0.00002f;
        L.s_shad = (128.0f - std::sqrt(std::fmax(lt->fog_offset, 0.0f)) * 11.313708f) * d50 * 0.28f;
        L.s_shad_zmul = d50 * 0.7f / (float)z_range * (128.0f - std::sqrt(std::fmax(lt->fog_far_offset, 0.0f)) * 11.313708f);
    }
    // rotate_object_lights(normalise_matrix_to(1.0, &sc.vgrads))
    double ms[3][3];
    for (int i = 0; i < 3; i++) {
        const double s = 1.0 / std::sqrt(sc->vgrads[i][0] * sc->vgrads[i][0] + sc->vgrads[i][1] * sc->vgrads[i][1] +
                                         sc->vgrads[i][2] * sc->vgrads[i][2]);
        for (int j = 0; j < 3; j++) ms[i][j] = sc->vgrads[i][j] * s;
    }
    for (int i = 0; i < n; i++) {
        mb3d_hip_light_val &o = L.lights[i];
        const mb3d_hip_light &li = lt->lights[o.idx];
        if (li.relative_to_object && !li.positional) {
            double r[3];
            for (int k = 0; k < 3; k++) r[k] = (double)o.ln[0] * ms[0][k] + (double)o.ln[1] * ms[1][k] + (double)o.ln[2] * ms[2][k];
            const double len = std::sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
            if (len > 0.0)
                for (int k = 0; k < 3; k++) r[k] *= 1.0 / len;
            for (int k = 0; k < 3; k++) o.ln[k] = (float)r[k];
        }
    }
    // place_lights(l, stereo_mid = mid, cam, z_range, view_z = ms[2])
    for (int i = 0; i < n; i++) {
        mb3d_hip_light_val &o = L.lights[i];
        const mb3d_hip_light &li = lt->lights[o.idx];
        if (o.positional)
            for (int k = 0; k < 3; k++) o.ln[k] = (float)(li.position[k] - sc->mid[k]);
        if (o.visible != 0) {
            const float z = calc_z_pos_for_light(cam, from_arr(o.ln), o.positional != 0, ms[2]);
            if (o.positional) {
                o.pos_z = z;
                const double v = 8388352.0 - cam.zc_mul * (std::sqrt((double)z / cam.step_width * cam.zcorr + 1.0) - 1.0);
                o.pos_lp = (float)clampd_rs(v / 256.0, 0.0, 32767.0);
            } else {
                o.pos_lp = 0.0f;
                o.pos_z = (float)z_range;
            }
        }
    }
    // SortLights (stable, by key)
    std::vector<mb3d_hip_light_val> v(L.lights, L.lights + n);
    std::stable_sort(v.begin(), v.end(), [](const mb3d_hip_light_val &a, const mb3d_hip_light_val &b) {
        const int ka = (a.visible & 6) != 0 ? (int)std::round(a.pos_lp) : 40000;
        const int kb = (b.visible & 6) != 0 ? (int)std::round(b.pos_lp) : 40000;
        return ka < kb;
    });
    for (int i = 0; i < n; i++) L.lights[i] = v[i];
    *cam_out = cam;
    *lv_out = L;
    return 0;
}

// gbuffer::pack (src/gbuffer.rs:64-80) of the card records read as
// march_card does (src/gpu.rs:3452-3463), and gbuffer::unpack
// (src/gbuffer.rs:83-98) written back as card records
void orc_gbuf_pack(const uint32_t *gbuf, int64_t n, uint8_t *out) {
    for (int64_t i = 0; i < n; i++) {
        const SiLight s = from_words(gbuf + 5 * i);
        uint8_t *d = out + 18 * i;
        auto le16 = [&](int at, uint16_t v) {
            d[at] = (uint8_t)(v & 0xFF);
            d[at + 1] = (uint8_t)(v >> 8);
        };
        for (int k = 0; k < 3; k++) le16(2 * k, (uint16_t)s.normal[k]);
        for (int k = 0; k < 4; k++) d[6 + k] = (uint8_t)(s.zpos_fine >> (8 * k));
        le16(10, s.shadow);
        le16(12, s.amb_shadow);
        le16(14, s.si_gradient);
        le16(16, s.otrap);
    }
}

void orc_gbuf_unpack(const uint8_t *in, int64_t n, uint32_t *gbuf) {
    for (int64_t i = 0; i < n; i++) {
        const uint8_t *b = in + 18 * i;
        auto h = [&](int k) { return (uint32_t)b[k] | ((uint32_t)b[k + 1] << 8); };
        const uint32_t zf = (uint32_t)b[6] | ((uint32_t)b[7] << 8) | ((uint32_t)b[8] << 16) | ((uint32_t)b[9] << 24);
        uint32_t *u = gbuf + 5 * i;
        u[0] = h(0) | (h(2) << 16);
        u[1] = h(4) | (h(10) << 16);
        u[2] = zf;
        u[3] = h(14) | (h(16) << 16);
        u[4] = h(12);
    }
}

float orc_cos_tab_val(int32_t tnr, float dotp, float rough, int32_t sqr) { return cos_tab_val(tnr, dotp, rough, sqr != 0); }

void orc_cos_tables(float *out) {
    const CosTabs &t = tabs();
    std::memcpy(out, t.diff, sizeof t.diff);
}

int32_t orc_paint(const mb3d_hip_paint_camera *cam, const mb3d_hip_light_vals *lv, const uint32_t *gbuf, int32_t x0,
                  int32_t y0, int32_t w, int32_t h, int32_t fmt, void *rgb, int32_t threads) {
    const int bpp = fmt == MB3D_HIP_PAINT_U8 ? 3 : (fmt == MB3D_HIP_PAINT_U16 ? 6 : 12);
    uint8_t *out = static_cast<uint8_t *>(rgb);
    const Lv lvw{*lv};
    (void)tabs();
    int nt = threads > 0 ? threads : (int)std::max(1u, std::thread::hardware_concurrency());
    nt = std::max(1, std::min(nt, h));
    auto rows = [&](int t) {
        for (int y = t; y < h; y += nt)
            for (int x = 0; x < w; x++) {
                const size_t i = (size_t)y * w + x;
                pixel_color(lvw, from_words(gbuf + 5 * i), x + x0, y + y0, *cam, fmt, out + bpp * i);
            }
    };
    std::vector<std::thread> pool;
    for (int t = 1; t < nt; t++) pool.emplace_back(rows, t);
    rows(0);
    for (auto &th : pool) th.join();
    return 0;
}

}  // extern "C"
