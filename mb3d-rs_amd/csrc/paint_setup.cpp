// paint_setup.cpp -- host side of the painter: the lighting defaults, the
// painter's camera, the derived light values and the diffuse look-up
// tables that the paint kernel reads.
//
//   mb3d_hip_lighting_default  Lighting::default        src/lighting.rs:290-392
//   mb3d_hip_paint_setup       paint_camera             src/render.rs:2340-2380
//                              light_vals               src/render.rs:2285-2319
//                              LightVals::build         src/lighting.rs:978-1252
//                              set_vol_light            src/lighting.rs:1273-1283
//                              rotate_object_lights     src/lighting.rs:1350-1381
//                              place_lights             src/lighting.rs:1285-1320
//   paint_cos_tables           cos_tabs                 src/lighting.rs:464-499
//
// The arithmetic keeps the reference's types: f32 where its LightVals
// fields are f32 (every product in its order), f64 where it computes in
// f64 and rounds at the end.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>

#include "../../include/mb3d_hip.h"
#include "host_common.h"
#include "paint_common.h"

namespace {

constexpr double kPi = 3.141592653589793;

// Delphi TColor ($BBGGRR) as [r, g, b] (rgb, src/lighting.rs:296-299)
void tcolor(uint32_t c, uint8_t out[3]) {
    out[0] = (uint8_t)(c & 0xFF);
    out[1] = (uint8_t)((c >> 8) & 0xFF);
    out[2] = (uint8_t)((c >> 16) & 0xFF);
}

void light_default(mb3d_hip_light &l) {
    std::memset(&l, 0, sizeof l);
    l.color[0] = l.color[1] = l.color[2] = 255;
    l.amplitude = 1.0f;
    l.spec_func = 4;
    l.hs_enabled = 1;
}

// round half away from zero, as Rust's f32::round
int round_away(float v) { return (int)std::round(v); }

// f32 `x.powi(n)` for n >= 1: multiplications by squaring from the low
// bit up, the order LLVM expands powi in (and compiler-rt's __powisf2)
float powi_f(float x, int n) {
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

struct Mat3 {
    double r[3][3];
};

Mat3 normalise_rows(const double m[3][3]) {  // normalise_matrix_to(1.0, m), src/math.rs:78-87
    Mat3 o;
    for (int i = 0; i < 3; i++) {
        const double d = 1.0 / std::sqrt(m[i][0] * m[i][0] + m[i][1] * m[i][1] + m[i][2] * m[i][2]);
        for (int j = 0; j < 3; j++) o.r[i][j] = m[i][j] * d;
    }
    return o;
}

void rotate_reverse(const double v[3], const Mat3 &m, double o[3]) {  // src/math.rs:63-69
    for (int k = 0; k < 3; k++) o[k] = v[0] * m.r[0][k] + v[1] * m.r[1][k] + v[2] * m.r[2][k];
}

// f32 vector helpers in the reference's operation order
inline float dot3(const float a[3], const float b[3]) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

// PaintCamera::view_vec (src/lighting.rs:748-772) on the host: f32 sin and
// cos, as the reference's f32 methods
void view_vec_host(const mb3d_hip_paint_camera &c, float x_pos, float y_pos, float aspect, float v[3]) {
    const float cx = (c.x_off - x_pos) * c.fov * aspect;
    const float cy = (y_pos - 0.5f) * c.fov;
    if (c.planar == 1) {
        float t[3] = {-cx, cy, c.pl_optic_z};
        const float n = std::sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
        for (int k = 0; k < 3; k++) v[k] = t[k] * (1.0f / n);
    } else if (c.planar == 2) {
        v[0] = -std::sin(cx) * std::cos(cy);
        v[1] = std::sin(cy);
        v[2] = std::cos(cy) * std::cos(cx);
    } else if (c.planar == 3) {
        const double xa = cy, ya = cx;  // build_view_vector_fisheye(cy, cx), src/math.rs:139-147
        const double t = std::sqrt(xa * xa + ya * ya);
        if (t < 1e-12) {
            v[0] = 0.0f, v[1] = 0.0f, v[2] = 1.0f;
        } else {
            const double tc = std::min(t, kPi);
            const double s = std::sin(tc) / t;
            v[0] = (float)(-ya * s), v[1] = (float)(xa * s), v[2] = (float)std::cos(tc);
        }
    } else {
        float t[3] = {-std::sin(cx), std::sin(cy), std::cos(cy) * std::cos(cx)};
        const float n = std::sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
        for (int k = 0; k < 3; k++) v[k] = t[k] * (1.0f / n);
    }
}

void to_abs_host(const mb3d_hip_paint_camera &c, const float v[3], float o[3]) {
    for (int k = 0; k < 3; k++) o[k] = v[0] * c.m[0][k] + v[1] * c.m[1][k] + v[2] * c.m[2][k];
}

double signum(double x) { return std::isnan(x) ? x : std::copysign(1.0, x); }

// CalcXYZposForLight, z only (calc_z_pos_for_light, src/lighting.rs:834-897)
float z_pos_for_light(const mb3d_hip_paint_camera &cam, const float lpos[3], bool positional,
                      const double view_z[3]) {
    if (!positional) return 0.0f;
    const float w = (float)cam.width, h = (float)cam.height;
    const float aspect = w / h;
    auto plv = [&](float mx, float my, float sp[3], float ab[3]) {
        for (int k = 0; k < 3; k++) sp[k] = (cam.start_pos[k] + cam.y_add[k] * my) + cam.x_add[k] * mx;
        float v[3];
        view_vec_host(cam, mx / w, my / h, aspect, v);
        to_abs_host(cam, v, ab);
    };
    auto dist = [&](const float ab[3], const float sp[3]) -> double {
        float v1[3], v2[3];
        for (int k = 0; k < 3; k++) v1[k] = lpos[k] - sp[k];
        v2[0] = v1[1] * (v1[2] - ab[2]) - v1[2] * (v1[1] - ab[1]);
        v2[1] = v1[2] * (v1[0] - ab[0]) - v1[0] * (v1[2] - ab[2]);
        v2[2] = v1[0] * (v1[1] - ab[1]) - v1[1] * (v1[0] - ab[0]);
        return (double)std::sqrt(dot3(v2, v2));
    };
    const double dmul = 0.001 / cam.step_width;
    float x = w * 0.5f, y = h * 0.5f;
    float sp[3], ab[3];
    plv(x, y, sp, ab);
    for (int it = 0; it <= 20; it++) {
        plv(x, y, sp, ab);
        const double d = dist(ab, sp) * dmul;
        if (d < 1e-4) break;
        float s1[3], a1[3], s2[3], a2[3];
        plv(x + 0.01f, y, s1, a1);
        const double dx = d - dist(a1, s1) * dmul;
        plv(x, y + 0.01f, s2, a2);
        const double dy = d - dist(a2, s2) * dmul;
        const double dd = std::sqrt(dx * dx + dy * dy) + 1e-16;
        std::memcpy(sp, s2, sizeof sp);
        std::memcpy(ab, a2, sizeof ab);
        if (dd * 2.0 > d) break;
        x += (float)(signum(dx) * d * 4e-3 / dd);
        y += (float)(signum(dy) * d * 4e-3 / dd);
    }
    float zf;
    if (std::fabs(ab[0]) > 0.5f)
        zf = (lpos[0] - sp[0]) / ab[0];
    else if (std::fabs(ab[1]) > 0.5f)
        zf = (lpos[1] - sp[1]) / ab[1];
    else
        zf = (lpos[2] - sp[2]) / ab[2];
    double z = zf;
    const double dz = cam.zz_stmit_dif;
    const double dv[3] = {(double)lpos[0] - view_z[0] * dz, (double)lpos[1] - view_z[1] * dz,
                          (double)lpos[2] - view_z[2] * dz};
    const double d = dv[0] * view_z[0] + dv[1] * view_z[1] + dv[2] * view_z[2];
    if (std::fabs(d) > std::fabs(z)) z = d;
    return (float)std::max(z, 0.0);
}

}  // namespace

namespace mb3d {

// cos_tabs (src/lighting.rs:464-499): 8 diffuse functions of 128 entries
// (the 4 light functions, then the 4 rough ones), f64 arithmetic rounded
// to f32 per entry.
void paint_cos_tables(float tab[8][128]) {
    for (int i = 0; i < 128; i++) {
        const double d = 1.0 - ((double)i - 2.0) / 60.0;
        if (d > 0.15)
            tab[0][i] = (float)((d - 0.08) * 1.0869565);
        else if (d <= 0.0)
            tab[0][i] = 0.0f;
        else
            tab[0][i] = (float)std::pow(d, std::max(1.0, (0.505 - d) * 3.8));
        const double dc = std::max(d, 0.0);
        tab[1][i] = (float)(dc * dc);
        tab[2][i] = (float)(d * 0.5 + 0.5);
        tab[3][i] = (float)((d * 0.5 + 0.5) * (d * 0.5 + 0.5));
    }
    for (int k = 0; k < 4; k++) {
        float tmp[128];
        for (int j = 0; j < 128; j++) tmp[j] = std::sqrt(std::max(tab[k][j], 0.0f));
        for (int j = 0; j < 128; j++) {
            double e = 0.0;
            for (int i = 0; i <= 60; i++) {
                const int l = std::abs(j + i - 30);
                if (l < 128) e += (double)tmp[l];
            }
            const double a = e * 0.011 + (e * 0.007) * (e * 0.007);
            tab[k + 4][j] = (float)(a * a);
        }
    }
}

}  // namespace mb3d

using namespace mb3d;

extern "C" void mb3d_hip_lighting_default(mb3d_hip_lighting *l) {
    if (!l) return;
    std::memset(l, 0, sizeof *l);
    static const uint32_t cols[9] = {5873889, 8837614, 0x8B491D, 2988346, 12248958, 0xFFC49F,
                                     11287584, 14579248, 7481121};
    for (int i = 0; i < 3; i++) {
        l->palette[i].position = (uint16_t)(i * 10922);
        tcolor(cols[i * 3], l->palette[i].diffuse);
        tcolor(cols[i * 3 + 1], l->palette[i].specular);
    }
    for (int i = 3; i < 10; i++) {
        l->palette[i] = l->palette[0];
        l->palette[i].position = (uint16_t)(32200 + (i - 3) * 84);
    }
    for (int i = 0; i < 3; i++) {
        l->interior_pos[i] = (uint16_t)(i * 10922);
        tcolor(cols[i * 3], l->interior_col[i]);
    }
    l->interior_pos[3] = 32700;
    std::memcpy(l->interior_col[3], l->interior_col[0], 3);
    for (auto &li : l->lights) light_default(li);
    // the main light: LFunction 53 (spec ^64, diffuse function 3)
    mb3d_hip_light &a = l->lights[0];
    a.on = 1;
    tcolor(0xA0E8FF, a.color);
    a.x_angle = 1500.0 * kPi / 16384.0;
    a.y_angle = 5200.0 * kPi / 16384.0;
    a.spec_func = 53 & 7;
    a.diff_func = (53 >> 4) & 3;
    mb3d_hip_light &b = l->lights[1];
    tcolor(2307911, b.color);
    b.x_angle = -2822.0 * kPi / 16384.0;
    b.y_angle = -7737.0 * kPi / 16384.0;
    b.spec_func = 16 & 7;
    b.diff_func = (16 >> 4) & 3;
    tcolor(0x8B491D, l->amb_top);
    tcolor(0xFFC49F, l->amb_bottom);
    tcolor(0x8B491D, l->depth_col);
    tcolor(0xFFC49F, l->depth_col2);
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
    for (auto &v : l->interior_spec) v = 160;
    l->fine_col_adj[0] = l->fine_col_adj[1] = 30;
    l->vol_light = -1;
}

extern "C" int32_t mb3d_hip_paint_setup(const mb3d_hip_scene *sc, const mb3d_hip_params *p,
                                        const mb3d_hip_lighting *l, mb3d_hip_paint_camera *cam_out,
                                        mb3d_hip_light_vals *lv_out) {
    if (!sc || !p || !l || !cam_out || !lv_out) return set_error(MB3D_HIP_ERR_INVALID, "NULL argument");
    for (int i = 0; i < MB3D_HIP_MAX_LIGHTS; i++)
        if (l->lights[i].map > 0) return set_error(MB3D_HIP_UNSUPPORTED, "light maps are not supported");
    if (l->bg_image) return set_error(MB3D_HIP_UNSUPPORTED, "background pictures are not supported");
    if (l->diff_map) return set_error(MB3D_HIP_UNSUPPORTED, "the diffuse colour map is not supported");
    if (l->bg_ambient) return set_error(MB3D_HIP_UNSUPPORTED, "the background as ambient light is not supported");
    if (sc->width <= 0 || sc->height <= 0 || p->width != sc->width || p->height != sc->height)
        return set_error(MB3D_HIP_ERR_INVALID, "scene and parameters differ in size");

    // ---- paint_camera (src/render.rs:2340-2380), common optic, no stereo
    mb3d_hip_paint_camera cam;
    std::memset(&cam, 0, sizeof cam);
    const int w = sc->width, h = sc->height;
    cam.width = w;
    cam.height = h;
    cam.planar = 0;
    cam.fov = (float)(sc->fov_y * (kPi / 180.0));
    const double dh = std::min(std::max((double)cam.fov * 0.5, 0.01), 1.5);
    const Mat3 m = normalise_rows(p->vgrads);
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) cam.m[i][j] = (float)m.r[i][j];
    const double sw = p->step_width;
    const double zz_stmit_dif = sc->z_start - sc->mid[2];
    const double sp0[3] = {-0.5 * (double)w * sw, -0.5 * (double)h * sw, zz_stmit_dif};
    double sp[3];
    rotate_reverse(sp0, m, sp);
    cam.aspect = (float)w / (float)h;
    cam.x_off = 0.5f;
    cam.pl_optic_z = (float)(std::cos(dh) * dh / std::sin(dh));
    cam.zcorr = p->zcorr;
    cam.zc_mul = p->zc_mul;
    cam.step_width = sw;
    cam.zz_stmit_dif = zz_stmit_dif;
    for (int k = 0; k < 3; k++) {
        cam.start_pos[k] = (float)sp[k];
        cam.x_add[k] = (float)(m.r[0][k] * sw);
        cam.y_add[k] = (float)(m.r[1][k] * sw);
    }

    // ---- LightVals::build (src/lighting.rs:978-1252), image_scale 1
    mb3d_hip_light_vals lv;
    std::memset(&lv, 0, sizeof lv);
    const bool sqr = l->internal_gamma2 != 0;
    const float light_scale = sqr ? 1.5f : 1.0f;
    auto lcol = [sqr](const uint8_t c[3], float o[3]) {  // RGBColToSVecNoScale(SQR)
        for (int k = 0; k < 3; k++) o[k] = sqr ? (float)c[k] * (float)c[k] / 255.0f : (float)c[k];
    };
    auto pcol = [sqr](const uint8_t c[3], float o[3]) {  // ColToSVec
        for (int k = 0; k < 3; k++)
            o[k] = sqr ? (float)c[k] * (float)c[k] * 0.0000153787f : (float)c[k] * (1.0f / 255.0f);
    };
    auto hs_calced = [l](int idx) { return ((l->hs_calced >> idx) & 1) != 0; };
    const double z_step_div = sc->z_step_div;
    const double z_range = sc->z_end - sc->z_start;
    const double step_width = 2.1345 / (sc->zoom * (double)w);  // Scene::step_width
    int n = 0;
    for (int idx = 0; idx < MB3D_HIP_MAX_LIGHTS; idx++) {
        const mb3d_hip_light &li = l->lights[idx];
        if (!li.on) continue;
        mb3d_hip_light_val &o = lv.lights[n++];
        const double dtmp = -li.x_angle, dtmp2 = li.y_angle;
        const double v[3] = {-std::sin(dtmp), std::sin(dtmp2), std::cos(dtmp2) * std::cos(dtmp)};
        const double nn = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        for (int k = 0; k < 3; k++) o.ln[k] = (float)(-v[k] / nn);
        float raw[3];
        lcol(li.color, raw);
        for (int k = 0; k < 3; k++) raw[k] = raw[k] * light_scale;
        float lmax;
        if (li.positional) {
            const float s = li.amplitude * 1.3f * light_scale;
            for (int k = 0; k < 3; k++) o.col[k] = raw[k] * s;
            lmax = 800.0f * ((o.col[0] + o.col[1]) / 3.0f + 128.0f * li.amplitude);
        } else {
            for (int k = 0; k < 3; k++) o.col[k] = raw[k] * li.amplitude;
            lmax = 800.0f * (raw[0] + raw[1] + raw[2] + 128.0f) * li.amplitude;
        }
        if (li.visible == 8) lmax *= 5.0f;
        o.idx = idx;
        o.on = 1;
        o.positional = li.positional ? 1 : 0;
        o.visible = li.visible;
        o.lmax_l = lmax;
        o.pow_func = 2 << (li.spec_func & 7);
        o.diff_func = hs_calced(idx) && l->hs_set_cos ? 0 : (li.diff_func & 3);
        o.sub_amb_sh = hs_calced(idx) != (li.hs_enabled != 0);
        o.hs_calced = hs_calced(idx) ? 1 : 0;
        o.hs_mask = l->hs_soft && hs_calced(idx) ? 0xFFFF : (0x400 << idx);
    }
    lv.n_lights = n;
    lv.s_depth = l->depth_fog * 0.8e-6f;
    float s_shad_gr = (l->dyn_fog - 53.0f) * 1.0f * (float)z_step_div * 0.00065f;
    float dt = 2.2f / (float)z_step_div;
    float dfmul = (float)z_step_div * 0.015f;
    float stmp = 128.0f;
    if (sc->dfog_on_it > 0) {
        dt *= 0.25f;
        s_shad_gr *= 4.0f;
        dfmul *= 4.0f;
    } else {
        stmp = 137.0f;
    }
    lv.s_shad_gr = s_shad_gr;
    lv.s_dyn_fog_mul = dfmul;
    lv.s_shad = (stmp - std::sqrt(std::max(l->fog_offset, 0.0f)) * 11.313708f) * dt * 0.28f;
    lv.s_shad_zmul =
        dt * 0.7f / (float)z_range * (128.0f - std::sqrt(std::max(l->fog_far_offset, 0.0f)) * 11.313708f);
    lv.s_amb_shad = l->amb_shadow / 53.0f;
    // CalcSCstartAndSCmul
    float s_c_start = powi_f((l->color_start + 30.0f) / 90.0f, 2) * 32767.0f - 10900.0f;
    float s_c_mul = powi_f((l->color_end + 30.0f) / 90.0f, 2) * 32767.0f - 10900.0f - s_c_start;
    if (l->fine_col_adj_on) {
        const float d = s_c_start + s_c_mul * ((float)l->fine_col_adj[1] - 30.0f) * 0.0166666666666666f;
        s_c_start += s_c_mul * ((float)l->fine_col_adj[0] - 30.0f) * 0.0166666666666666f;
        s_c_mul = d - s_c_start;
    }
    if (std::fabs(s_c_mul) > 0.001f)
        s_c_mul = 2.0f / s_c_mul;
    else
        s_c_mul = s_c_mul < 0.0f ? -2000.0f : 2000.0f;
    lv.s_c_start = s_c_start;
    lv.s_c_mul = s_c_mul;
    const float dti = powi_f(l->interior_start, 4);
    lv.s_ci_start = dti * 0.000158025f + 32768.0f;
    float s_ci_mul = (powi_f(l->interior_end, 4) - dti) * 0.5f * 0.00031605f;
    if (std::fabs(s_ci_mul) > 1e-10f) s_ci_mul = 1.0f / s_ci_mul;
    lv.s_ci_mul = s_ci_mul;
    const int gi = round_away(l->gamma);
    if (gi == 32) {
        lv.gamma_h = 0, lv.s_gamma = 0.0f;
    } else if (gi < 32) {
        lv.gamma_h = -1, lv.s_gamma = 1.0f - (float)gi / 32.0f;
    } else {
        lv.gamma_h = 1, lv.s_gamma = (float)(gi - 32) / 31.0f;
    }
    float amb_mul = l->ambient / 90.0f;
    if (sqr) amb_mul = (amb_mul * amb_mul + amb_mul) * 0.5f * light_scale;
    for (int i = 0; i < 10; i++) {
        pcol(l->palette[i].diffuse, lv.col_dif[i]);
        pcol(l->palette[i].specular, lv.col_spe[i]);
        const uint8_t *s = l->palette[i].specular;
        const float a = (float)(l->palette_alpha_set ? l->palette_alpha[i] : std::max(std::max(s[0], s[1]), s[2]));
        lv.col_spe_a[i] = sqr ? a * a * 0.0000153787f : a / 255.0f;
        lv.col_pos[i] = l->palette[i].position;
    }
    for (int i = 0; i < 10; i++) {
        const int j = i < 9 ? lv.col_pos[i + 1] - lv.col_pos[i] : 32767 - lv.col_pos[9];
        lv.s_c_div[i] = j > 1 ? 1.0f / (float)j : 1.0f;
    }
    for (int i = 0; i < 4; i++) {
        pcol(l->interior_col[i], lv.col_int[i]);
        lv.icol_pos[i] = l->interior_pos[i];
    }
    for (int i = 0; i < 4; i++) {
        const int j = i < 3 ? lv.icol_pos[i + 1] - lv.icol_pos[i] : 32767 - lv.icol_pos[3];
        lv.s_ic_div[i] = j > 1 ? 1.0f / (float)j : 1.0f;
    }
    const float ind = (float)round_away(std::min(std::max(l->ind_light - 53.0f, -128.0f), 127.0f));
    float s_ind = powi_f((ind + 53.0f) * 0.022f, 2);
    float s_diff = l->diffuse * 0.02f;
    float s_spec = std::max(l->specular * 0.02f, 0.004f);
    if (sqr) {
        s_ind *= 0.5f;
        s_diff = (s_diff * s_diff + s_diff) * 0.5f;
        s_spec = (s_spec * s_spec + s_spec) * 0.5f;
    }
    lv.s_ind_light_reflect = s_ind;
    lv.s_diff = s_diff;
    lv.s_spec = s_spec;
    for (int i = 0; i < 4; i++) {
        const float a = (float)l->interior_spec[i];
        lv.col_int_spec[i] = sqr ? a * a * 0.0000153787f : a / 255.0f;
    }
    float t[3];
    lcol(l->amb_top, t);
    for (int k = 0; k < 3; k++) lv.amb_col[k] = t[k] * amb_mul;
    lcol(l->amb_bottom, t);
    for (int k = 0; k < 3; k++) lv.amb_col2[k] = t[k] * amb_mul;
    lcol(l->depth_col, lv.depth_col);
    lcol(l->depth_col2, lv.depth_col2);
    lcol(l->dyn_fog_col, lv.dyn_fog_col);
    lcol(l->dyn_fog_col2, lv.dyn_fog_col2);
    lv.s_col_zmul = (float)((double)l->var_col_z * -0.005 / (step_width * (double)w));
    lv.s_roughness_factor = l->roughness / (255.0f * 255.0f);
    lv.col_cycling = l->color_cycling ? 1 : 0;
    lv.col_on_otrap = l->color_on_otrap ? 1 : 0;
    lv.far_fog = l->far_fog ? 1 : 0;
    lv.depth_func = l->depth_func;
    lv.s_diffuse_shadowing = l->diffuse_shadowing;
    lv.sqr = sqr ? 1 : 0;
    lv.no_col_ipol = l->no_col_ipol ? 1 : 0;
    lv.amb_rel_obj = l->amb_rel_obj ? 1 : 0;
    lv.has_amb_frame = l->has_amb_frame ? 1 : 0;
    if (l->has_amb_frame) {
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) lv.amb_frame[i][j] = (float)l->amb_frame[i][j];
        lv.amb_fov = (float)l->amb_fov;
    }
    lv.dfog_options = (l->dfog_options & 3) == 3 ? 1 : (l->dfog_options & 3);
    lv.ex_mode = l->ex_mode != 0 ? 1 : 0;

    // ---- set_vol_light (light_vals: the volumetric light's light is on)
    if (l->vol_light >= 0 && l->vol_light < MB3D_HIP_MAX_LIGHTS && l->lights[l->vol_light].on) {
        lv.vol_light = 1;
        lv.s_dyn_fog_mul = 0.0005f;
        const float d50 = 50.0f;
        lv.s_shad_gr = (l->dyn_fog - 53.0f) * 0.00002f;
        lv.s_shad = (128.0f - std::sqrt(std::max(l->fog_offset, 0.0f)) * 11.313708f) * d50 * 0.28f;
        lv.s_shad_zmul =
            d50 * 0.7f / (float)z_range * (128.0f - std::sqrt(std::max(l->fog_far_offset, 0.0f)) * 11.313708f);
    }

    // ---- rotate_object_lights with the scene's unit view matrix
    const Mat3 ms = normalise_rows(sc->vgrads);
    for (int i = 0; i < n; i++) {
        mb3d_hip_light_val &o = lv.lights[i];
        const mb3d_hip_light &li = l->lights[o.idx];
        if (li.relative_to_object && !li.positional) {
            const double v[3] = {o.ln[0], o.ln[1], o.ln[2]};
            double r[3];
            rotate_reverse(v, ms, r);
            const double len = std::sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
            if (len > 0.0)
                for (int k = 0; k < 3; k++) r[k] = r[k] * (1.0 / len);
            for (int k = 0; k < 3; k++) o.ln[k] = (float)r[k];
        }
    }

    // ---- place_lights (mid = the scene middle without stereo)
    const double *view_z = ms.r[2];
    for (int i = 0; i < n; i++) {
        mb3d_hip_light_val &o = lv.lights[i];
        const mb3d_hip_light &li = l->lights[o.idx];
        if (o.positional)
            for (int k = 0; k < 3; k++) o.ln[k] = (float)(li.position[k] - sc->mid[k]);
        if (o.visible != 0) {
            const float z = z_pos_for_light(cam, o.ln, o.positional != 0, view_z);
            if (o.positional) {
                o.pos_z = z;
                const double v = 8388352.0 - cam.zc_mul * (std::sqrt((double)z / cam.step_width * cam.zcorr + 1.0) - 1.0);
                o.pos_lp = (float)std::min(std::max(v / 256.0, 0.0), 32767.0);
            } else {
                o.pos_lp = 0.0f;
                o.pos_z = (float)z_range;
            }
        }
    }
    // SortLights: visible lights back to front by depth, the others after them (stable)
    auto key = [](const mb3d_hip_light_val &o) { return (o.visible & 6) != 0 ? round_away(o.pos_lp) : 40000; };
    std::stable_sort(lv.lights, lv.lights + n,
                     [&](const mb3d_hip_light_val &a, const mb3d_hip_light_val &b) { return key(a) < key(b); });
    *cam_out = cam;
    *lv_out = lv;
    return MB3D_HIP_OK;
}
