// mb3d_oracle.cpp -- TEST INFRASTRUCTURE (the f64 CPU oracle), not product code.
//
// A restatement of mb3d-rs's CPU ray marcher for the built-in formulas
// and their alternating 3D and 4D hybrids, in the reference's operation
// order and with its f32 roundings kept (so that the results are those of
// the reference's own CPU path):
//   parameter set-up   CalcParams::new          src/calc.rs:375-898
//   iteration          Iteration::run_with ...  src/iteration.rs:140-154,268-315,521-585,757-772
//   formulas           Formula::iterate         src/formulas.rs:576-858
//   4D rotation        build_smatrix4, rotate_4dex  src/math.rs:178-215
//   distance estimate  calc_de_part             src/calc.rs:1495-1559
//   march              march_pixel_from, march_from  src/calc.rs:2231-2322,2533-2657
//   bisection          bin_search, bin_search_it     src/calc.rs:1706-1792
//   normals            calculate_normals, smoothing sweeps, calc_roughness,
//                      calculate_normals_on_smooth_it   src/calc.rs:1794-2006
//   surface            surface_found, do_color, do_color_on_it,
//                      calc_zpos_and_rough  src/calc.rs:2010-2090,2194-2205,2471-2517
//   cone march         cone_march, Grid::safe_depth (one level of blocks)
//                      src/calc.rs:2357-2435, src/cone.rs:66-115
//   hard shadows       hard_shadow_row, per pixel (no row shortcut)  src/calc.rs:2741-2899
//   DE AO              deao_row with dithering and the random first step,
//                      surface_point   src/deao.rs:254-418, src/calc.rs:2157-2192
//   secondary rays     SrTracer::march_ray, calc_hs, deao_1pos  src/reflect.rs:805-1090,1464-1674
//   Monte Carlo        McMarch::run (ray, shadow, through)      src/mc.rs:488-977
// Built with -ffp-contract=off: no fused multiply-adds, like the reference.
//
// Every function here is checked against the reference's known answers,
// closed forms or brute force in tests/.  The DE AO (deao_row) per pixel:
// its surface point against the march's depth, its ray directions
// (dithered too) against a Rodrigues rotation of the ray set onto the
// normal, its openness against a fine DE walk along each ray from the
// (random) first step (test_deao_against_brute_force); each walk's
// samples against the DE evaluated apart at their points, each step the
// larger of that DE and the walk's growth, the end at the openness floor or
// the ray's length
// (test_deao_walk_steps_by_the_larger_of_the_de_and_its_growth); its
// generator against the C runtime's rand() outputs and the row's draw
// order (test_deao_random_first_steps_draw_along_the_row); the hit pass's
// deao_1pos against deao_row on view rays (test_oracle_rays.py).  The DE's
// size against the brute-force distance to the set, analytic and numeric
// (test_oracle.py).  The smoothing sweeps and normals on the smooth
// iteration against the DE gradient (test_oracle_normals.py); the cone
// march against the reference's own cone tests (test_oracle_cone.py).
// Colour options 1..5 and colour on an iteration (do_color,
// do_color_on_it): the orbit end they read against an independent iterate
// of the surface point, each word inverted back to that orbit end
// (test_oracle_normals.py).
#include "mb3d_oracle.h"

#include <algorithm>
#include <array>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

namespace {

thread_local std::string g_err;

int fail(const char *msg) {
    g_err = msg;
    return MB3D_HIP_ERR_INVALID;
}

typedef double Vec3[3];
typedef double Mat3[3][3];

// math.rs helpers (src/math.rs:26-133)
inline void add_weight(double a[3], const double b[3], double w) {
    a[0] += b[0] * w;
    a[1] += b[1] * w;
    a[2] += b[2] * w;
}
inline double dot3(const double a[3], const double b[3]) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

// rotate_vector_reverse (src/math.rs:63-69): V' = V * M
inline void rot_rev(const double v[3], const double m[3][3], double r[3]) {
    r[0] = v[0] * m[0][0] + v[1] * m[1][0] + v[2] * m[2][0];
    r[1] = v[0] * m[0][1] + v[1] * m[1][1] + v[2] * m[2][1];
    r[2] = v[0] * m[0][2] + v[1] * m[1][2] + v[2] * m[2][2];
}

// normalise_matrix_to (src/math.rs:78-87)
void normalise_matrix_to(double n, const double m[3][3], double r[3][3]) {
    for (int i = 0; i < 3; i++) {
        double d = n / std::sqrt(dot3(m[i], m[i]));
        for (int j = 0; j < 3; j++) r[i][j] = m[i][j] * d;
    }
}

// build_view_vector_dfov (src/math.rs:127-133)
void build_view_vector_dfov(double xa, double ya, double v[3]) {
    double sx = std::sin(xa), cx = std::cos(xa);
    double sy = std::sin(ya), cy = std::cos(ya);
    double w[3] = {-sy, sx, cx * cy};
    double n = 1.0 / std::sqrt(dot3(w, w));
    v[0] = w[0] * n;
    v[1] = w[1] * n;
    v[2] = w[2] * n;
}

// min_max_clip_15bit (src/math.rs:246-255)
uint16_t clip15(float s) {
    if (std::isnan(s) || s < 0.0f) return 0;
    if (s > 32767.0f) return 32767;
    return (uint16_t)std::nearbyint(s);  // round half to even (default rounding mode)
}

inline double round_even(double x) { return std::nearbyint(x); }

// ---------------------------------------------------------------------------
// formulas (src/formulas.rs)

// box_fold (src/formulas.rs:707-710)
inline double box_fold(double x, double fold) { return std::fabs(x + fold) - (std::fabs(x - fold) + x); }

// amazing_box (src/formulas.rs:713-736)
inline void amazing_box(double scale, double min_r, double fold, double v[4], const double j[4], bool ade) {
    double x = box_fold(v[0], fold);
    double y = box_fold(v[1], fold);
    double z = box_fold(v[2], fold);
    double r = z * z + y * y + x * x;
    double m = std::fmax(min_r, 1e-40);
    double sqr_min_r = m * m;
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

// int_pow2_scale (src/formulas.rs:739-752): HybridItIntPow2scale of Bulbox
inline void int_pow2_scale(double scaling, double v[4], const double j[4]) {
    double s = 1.0 / scaling;
    double x = v[0] * s, y = v[1] * s, z = v[2] * s;
    double xx = x * x, yy = y * y, r = xx + yy;
    v[2] = -((std::sqrt(r) * z) * 2.0) * scaling + j[2];
    double a = (r - z * z) / r;
    v[0] = (xx - yy) * a * scaling + j[0];
    v[1] = (x * y * a) * 2.0 * scaling + j[1];
}

// real_power (src/formulas.rs:755-766): HybridFloatPow, radius from the
// stored squared radius of the previous counted iteration
inline void real_power(double pw, double z_mul, double v[4], const double j[4], double rout) {
    double x = v[0], y = v[1], z = v[2];
    double theta = std::atan2(y, x) * pw;
    double sin_t = std::sin(theta), cos_t = std::cos(theta);  // f64::sin_cos
    double phi = std::atan2(z, std::sqrt(x * x + y * y)) * pw;
    double sin_p = std::sin(phi), cos_p = std::cos(phi);
    double r = std::pow(rout, pw * 0.5);
    v[0] = cos_t * cos_p * r + j[0];
    v[1] = sin_t * cos_p * r + j[1];
    v[2] = sin_p * r * z_mul + j[2];
}

// Quaternion (src/formulas.rs:580-586)
inline void quaternion(double yw_mul, double w_add, double v[4], const double j[4]) {
    double x = v[0], y = v[1], z = v[2], w = v[3];
    v[0] = x * x - y * y - z * z - w * w + j[0];
    v[1] = 2.0 * (y * x + z * w) + j[1];
    v[2] = 2.0 * (z * x + yw_mul * y * w) + j[2];
    v[3] = 2.0 * (w * x + y * z) + w_add + j[3];
}

// Tricorn (src/formulas.rs:587-592)
inline void tricorn(double z_mul, double cz_mul, double v[4], const double j[4]) {
    double x = v[0], y = v[1], z = v[2];
    v[2] = z * x * z_mul + j[2] * cz_mul;
    v[0] = x * x - (y * y + z * z) + j[0];
    v[1] = 2.0 * (y * x) + j[1];
}

// Bulbox, HybridSuperCube2 (src/formulas.rs:596-627): the power-2 bulb
// inside R threshold 2, the box outside R threshold, blended between
inline void bulbox(const double o[8], double v[4], const double j[4], double rout) {
    double scale = o[0], min_r = o[1], fold = o[2], bulb_scale = o[3];
    double rt1 = o[4] * o[4];
    double rt2 = o[5] * o[5];
    if (rout < rt1) {
        if (rout < rt2) {
            int_pow2_scale(bulb_scale, v, j);
        } else {
            double m = (rout - rt2) / (rt1 - rt2);
            double in[3] = {v[0], v[1], v[2]};
            int_pow2_scale(bulb_scale, v, j);
            double out[3] = {v[0], v[1], v[2]};
            v[0] = in[0];
            v[1] = in[1];
            v[2] = in[2];
            amazing_box(scale, min_r, fold, v, j, false);
            double r1 = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
            double m2 = 1.0 - m;
            double r2 = r1 * m + m2 * std::sqrt(out[0] * out[0] + out[1] * out[1] + out[2] * out[2]);
            v[0] = v[0] * m + m2 * out[0];
            v[1] = v[1] * m + m2 * out[1];
            v[2] = v[2] * m + m2 * out[2];
            r1 = r2 / std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2] + 1e-40);
            v[0] *= r1;
            v[1] *= r1;
            v[2] *= r1;
        }
    } else {
        amazing_box(scale, min_r, fold, v, j, false);
    }
}

// AexionC (src/formulas.rs:644-703): a real power bulb step with the angles
// from the y axis, then (rotate C) the constant turned by a power of its
// own angles.  o: power, z_mul, rotate C, conditional phi, power C, cz_mul,
// power C on |Z - C|, mode.
inline void aexion_c(const double o[8], double v[4], double j[4]) {
    double power = o[0], z_mul = o[1], power_c = o[4], cz_mul = o[5];
    bool rot_c = o[2] != 0.0, cond_phi = o[3] != 0.0, powc_dist = o[6] != 0.0;
    int mode = (int)o[7];
    double x = v[0], y = v[1], z = v[2];
    double r1 = x * x + y * y + z * z;
    double th = std::atan2(std::sqrt(x * x + z * z), y) * power;
    double ph = std::atan2(z, x) * power;
    r1 = std::exp(std::log(r1) * power * 0.5);
    double st = std::sin(th), ct = std::cos(th);
    double sp = std::sin(ph), cp = std::cos(ph);
    v[0] = cp * ct * r1 + j[0];
    v[2] = st * r1 * z_mul + j[2];
    v[1] = ct * sp * r1 + j[1];
    if (!rot_c) return;
    double pd = power_c;
    if (powc_dist) {
        double d0 = v[0] - j[0], d1 = v[1] - j[1], d2 = v[2] - j[2];
        pd *= std::sqrt(d0 * d0 + d1 * d1 + d2 * d2);
    }
    double r = std::sqrt(j[0] * j[0] + j[1] * j[1] + j[2] * j[2]);
    double a, b, c;
    if (mode & 16) {
        a = v[1] - j[1];
        b = v[2] - j[2];
        c = v[0] - j[0];
    } else {
        a = j[1];
        b = j[2];
        c = j[0];
    }
    int ofs = 0, ecx = 8;  // the component "conditional phi" tests: 0 x, 8 y, 16 z
    if (mode & 8) {
        std::swap(a, b);
        ecx = 16;
    }
    double s = std::sqrt(b * b + c * c);
    double t1 = (mode & 1) ? std::atan2(a, s) : std::atan2(s, a);
    double p1;
    if (mode & 2) {
        ofs = 24 - ecx;
        p1 = std::atan2(c, b);
    } else {
        p1 = std::atan2(b, c);
    }
    if (mode & 4) {
        std::swap(t1, p1);
        ofs = ecx;
    }
    if (cond_phi) {
        double t = ofs == 0 ? v[0] : (ofs == 8 ? v[1] : v[2]);
        if (!std::signbit(t)) p1 = -p1;
    }
    double sa = std::sin(p1 * pd), ca = std::cos(p1 * pd);
    double sb = std::sin(t1 * pd), cb = std::cos(t1 * pd);
    j[0] = ca * cb * r;
    j[2] = sb * r * cz_mul;
    j[1] = cb * sa * r;
}

// int_pow (src/formulas.rs:769-858)
void int_pow(int power, double zmul, double v[4], const double j[4]) {
    double x = v[0], y = v[1], z = v[2];
    const double EPS = 1e-40;
    if (power <= 2) {
        double yy = y * y, xx = x * x, r = xx + yy;
        v[2] = (std::sqrt(r) * zmul * z) * 2.0 + j[2];
        double a = (r - z * z) / r;
        v[0] = (xx - yy) * a + j[0];
        v[1] = (x * y * a) * 2.0 + j[1];
    } else if (power == 3) {
        double sy = y * y, sx = x * x, r = sx + sy, sz = z * z;
        double a = 1.0 - 3.0 * sz / (r + EPS);
        v[0] = (sx - 3.0 * sy) * a * x + j[0];
        v[2] = j[2] - (sz - 3.0 * r) * z * zmul;
        v[1] = a * (3.0 * sx - sy) * y + j[1];
    } else if (power == 4) {
        double sy = y * y, sx = x * x, r = sx + sy, sz = z * z;
        double a = 1.0 + sz * (sz - 6.0 * r) / (r * r + EPS);
        v[0] = (sx * (sx - 6.0 * sy) + sy * sy) * a + j[0];
        v[2] = std::sqrt(r) * (r - sz) * z * 4.0 * zmul + j[2];
        v[1] = a * x * (sx - sy) * 4.0 * y + j[1];
    } else if (power == 5) {
        double sy = y * y, sx = x * x, r = sx + sy, sz = z * z;
        double a = 1.0 + (sz * sz - sz * r * 2.0) * 5.0 / (r * r + EPS);
        v[1] = (5.0 * sx * sx - sy * (10.0 * sx - sy)) * a * y + j[1];
        v[2] = (sz * (sz - 10.0 * r) + 5.0 * r * r) * z * zmul + j[2];
        v[0] = a * x * (sx * (sx - 10.0 * sy) + 5.0 * sy * sy) + j[0];
    } else if (power == 6) {
        double sy = y * y, sx = x * x, r = sx + sy, sz = z * z, rr = r * r;
        double a = 1.0 - sz * (15.0 * rr + sz * (sz - r * 15.0)) / (rr * r + EPS);
        v[1] = (3.0 * sy * sy + sx * (3.0 * sx - 10.0 * sy)) * a * y * x * 2.0 + j[1];
        v[2] = (sz * (3.0 * sz - 10.0 * r) + 3.0 * rr) * std::sqrt(r) * z * zmul * 2.0 + j[2];
        v[0] = (sy * sy * (15.0 * sx - sy) + sx * sx * (sx - 15.0 * sy)) * a + j[0];
    } else if (power == 7) {
        double sy = y * y, sx = x * x, r = sx + sy, sz = z * z, rr = r * r;
        double a = 1.0 - 7.0 * (sz * (3.0 * rr + sz * (sz - 5.0 * r))) / (rr * r + EPS);
        v[1] = (sx * (21.0 * sy * sy + sx * (7.0 * sx - 35.0 * sy)) - sy * sy * sy) * a * y + j[1];
        v[2] = j[2] - (sz * sz * sz - 7.0 * r * (rr + sz * (3.0 * sz - 5.0 * r))) * z * zmul;
        v[0] = a * x * (sx * (35.0 * sy * sy + sx * (sx - 21.0 * sy)) - 7.0 * sy * sy * sy) + j[0];
    } else {
        // HybridIntP8 (White's / Nylander's power 8)
        double xx = x * x, yy = y * y, zz = z * z;
        double r = xx + yy;
        double rr = r * r;
        double zzzz = zz * zz;
        double t = (zzzz - 6.0 * r * zz + rr) * (zz - r) * std::sqrt(r);
        v[2] = -(t * z * 8.0 * zmul) + j[2];
        double num = (rr * 70.0 + zzzz) * zzzz - 28.0 * zz * r * (zzzz + rr);
        double a = num / (rr * rr + EPS) + 1.0;
        double xxxx = xx * xx, yyyy = yy * yy;
        v[1] = (yyyy * (7.0 * xx - yy) + xxxx * (xx - 7.0 * yy)) * 8.0 * x * y * a + j[1];
        v[0] = a * (xxxx * xxxx + yyyy * (70.0 * xxxx + yyyy) - 28.0 * (xx * yy * (yyyy + xxxx))) + j[0];
    }
}

// ---------------------------------------------------------------------------
// the iteration record and the hybrid loops (src/iteration.rs)

struct Iter {
    double v[4] = {0, 0, 0, 0};
    double c[3] = {0, 0, 0};
    double j[4] = {0, 0, 0, 0};
    double rout = 0, rold = 0, otrap = 0;
    int it_result = 0, max_it = 0;
    double rstop = 0;
    bool calc_sit = false;
    float smooth_it = 0;
    float ln_rstop = 0;
    int de_option = 0;
    double deriv1 = 1.0;
    int start_from = 0, end_to = 0, repeat_from = 0;
    float fhln[6] = {0, 0, 0, 0, 0, 0};
    bool four_d = false, do_julia = false, ipol = false;
    double ju[4] = {0, 0, 0, 0};
    float ipw[2] = {1.0f, 0.0f};
    uint64_t its_p8 = 0, its_box = 0, its_other = 0;
};

struct Marcher {
    const mb3d_hip_params &p;
    Iter it;
    float ms_de_stop;
    double mzz = 0;
    double vfov[3] = {0, 0, 1};
    double cafy = 0;
    int max_its_result;
    // the hybrid parts (HybridPart, src/calc.rs:137-150): the picture's one
    // part, or the two of a DE combination
    struct PartO {
        int start, end, repeat, max_it, de_option;
        bool custom, four;
        float dds;
    };
    PartO part1{}, part2{}, part{};
    int32_t seed = 0;
    uint64_t de_evals = 0;
    int nrm_it = 0;  // the iteration count of calculate_normals' centre DE
    float s_roughness = 0.0f;  // calc_roughness of the last smoothed normal
    bool calc_inside = false;  // the DE from inside the set (the Monte Carlo marches' inside flag)

    explicit Marcher(const mb3d_hip_params &pp) : p(pp) {
        // CalcParams::new_iteration (src/calc.rs:1054-1087)
        it.max_it = p.max_it;
        it.rstop = p.d_rstop;
        it.ln_rstop = p.ln_rstop;
        it.de_option = p.de_option;
        it.deriv1 = 1.0;
        it.end_to = p.end_to;
        it.repeat_from = p.repeat_from;
        for (int k = 0; k < 6; k++) it.fhln[k] = p.fhln[k];
        it.four_d = p.four_d != 0;
        it.do_julia = p.do_julia != 0;
        it.ipol = p.ipol != 0;
        it.ipw[0] = p.ipol_w[0];
        it.ipw[1] = p.ipol_w[1];
        for (int k = 0; k < 4; k++) it.ju[k] = p.julia_c[k];
        for (int k = 0; k < 4; k++) it.j[k] = p.julia_c[k];  // j starts as ju (src/calc.rs:1058)
        ms_de_stop = p.de_stop;
        max_its_result = p.max_it;
        part1 = {0, p.end_to, p.repeat_from, p.max_it, p.de_option, p.is_custom_de != 0, p.four_d != 0, p.d_de_scale};
        part2 = {p.start2, p.end2, p.repeat2, p.max_it2, p.de_option2, p.is_custom_de2 != 0, p.four_d2 != 0,
                 p.d_de_scale2};
        part = part1;
    }

    // Formula::iterate through call_slot (src/iteration.rs:325-345,
    // src/formulas.rs:576-641); the analytic power-8 extension keeps the
    // running derivative dr <- p r^(p-1) dr + 1 in deriv1.
    void call_slot(int n) {
        int kind = p.slot_kind[n];
        if (kind == MB3D_HIP_FORMULA_INTPOW) {
            if (p.slot_ade[n]) {
                double r2 = it.v[0] * it.v[0] + it.v[1] * it.v[1] + it.v[2] * it.v[2];
                int pw = p.slot_power[n];
                double rp = pw == 8 ? r2 * r2 * r2 * std::sqrt(r2) : std::pow(r2, 0.5 * (pw - 1));
                it.deriv1 = (double)pw * rp * it.deriv1 + 1.0;
            }
            int_pow(p.slot_power[n], p.slot_z_mul[n], it.v, it.j);
            it.its_p8++;
        } else if (kind == MB3D_HIP_FORMULA_AMAZINGBOX) {
            amazing_box(p.slot_scale[n], p.slot_min_r[n], p.slot_fold[n], it.v, it.j, p.slot_ade[n] != 0);
            it.its_box++;
        } else {
            const double *o = p.slot_opt[n];
            switch (kind) {
                case MB3D_HIP_FORMULA_REALPOWER: real_power(o[0], o[1], it.v, it.j, it.rout); break;
                case MB3D_HIP_FORMULA_QUATERNION: quaternion(o[0], o[1], it.v, it.j); break;
                case MB3D_HIP_FORMULA_TRICORN: tricorn(o[0], o[1], it.v, it.j); break;
                case MB3D_HIP_FORMULA_BULBOX: bulbox(o, it.v, it.j, it.rout); break;
                case MB3D_HIP_FORMULA_FOLDINGINTPOW:  // src/formulas.rs:629-634
                    for (int k = 0; k < 3; k++) it.v[k] = box_fold(it.v[k], o[2]);
                    int_pow(p.slot_power[n], o[1], it.v, it.j);
                    break;
                case MB3D_HIP_FORMULA_AEXIONC: aexion_c(o, it.v, it.j); break;
                default: break;
            }
            it.its_other++;
        }
    }

    // Iteration::calc_smooth_iterations (src/iteration.rs:140-154)
    void calc_smooth_iterations(int n) {
        if (it.rout <= 1.0) {
            it.smooth_it = (float)it.it_result;
        } else if (it.rold < 1.0) {
            double d = std::log(std::log(it.rout) * 0.5) * (double)it.fhln[n];
            it.smooth_it = (float)((double)it.it_result + (double)it.ln_rstop - d);
        } else {
            double d = std::log(0.5 * std::log(it.rout));
            double e = d - std::log(0.5 * std::log(it.rold));
            it.smooth_it = (float)(((double)it.ln_rstop - d) / (e + 1e-100) + (double)it.it_result);
        }
    }

    // Iteration::run_with (src/iteration.rs:268-315), 3D
    int run() {
        int n = it.start_from;
        int btmp = std::max(p.slot_iters[n], 0);
        it.it_result = 0;
        uint32_t uncounted_guard = 0;
        for (;;) {
            it.rold = it.rout;
            int guard = 0;
            while (btmp <= 0) {
                n += 1;
                if (n > it.end_to) n = it.repeat_from;
                btmp = std::max(p.slot_iters[n], 0);
                guard += 1;
                if (guard > 64) return n;
            }
            call_slot(n);
            btmp -= 1;
            if (p.slot_uncounted[n]) {
                uncounted_guard += 1;
                if (uncounted_guard > 100000) return n;
                continue;
            }
            it.it_result += 1;
            double x = it.v[0], y = it.v[1], z = it.v[2], w = it.v[3];
            it.rout = it.four_d ? x * x + y * y + z * z + w * w : x * x + y * y + z * z;
            if (it.rout < it.otrap) it.otrap = it.rout;
            // NaN compares false: keeps iterating
            if (it.it_result >= it.max_it || it.rout > it.rstop) return n;
        }
    }

    // the constant of a 3D hybrid: the point, or the Julia seed; j[3] (J4)
    // is kept (src/iteration.rs:523-527)
    void start_j3() {
        const double *c = it.do_julia ? it.ju : it.c;
        it.j[0] = c[0];
        it.j[1] = c[1];
        it.j[2] = c[2];
    }

    // Iteration::hybrid_4d (src/iteration.rs:757-772) with start_4d
    // (src/iteration.rs:417-428): the point rotated into 4D
    void start_4d() {
        const double *c = it.c;
        double v4[4];
        for (int i = 0; i < 4; i++)  // rotate_4dex (src/math.rs:209-215)
            v4[i] = c[0] * p.smatrix4[i][0] + c[1] * p.smatrix4[i][1] + c[2] * p.smatrix4[i][2];
        for (int k = 0; k < 4; k++) {
            it.v[k] = v4[k];
            it.j[k] = it.do_julia ? it.ju[k] : v4[k];
        }
        it.rout = v4[0] * v4[0] + v4[1] * v4[1] + v4[2] * v4[2] + v4[3] * v4[3];
        it.otrap = it.rout;
    }

    void hybrid_4d() {
        start_4d();
        int n = run();
        if (it.calc_sit) calc_smooth_iterations(n);
    }

    // Iteration::run_ipol (src/iteration.rs:347-401): both formulas on the
    // same vector, their results and lengths blended by the weights; `de`
    // blends the running derivative of the analytic DE in w (3D)
    void run_ipol(bool de) {
        const double s1 = (double)it.ipw[0], s2 = (double)it.ipw[1];
        const bool four = it.four_d;
        it.it_result = 0;
        for (;;) {
            it.rold = it.rout;
            double y1[4] = {it.v[0], it.v[1], it.v[2], it.v[3]};
            call_slot(0);
            double x1[4] = {it.v[0], it.v[1], it.v[2], it.v[3]};
            for (int k = 0; k < 4; k++) it.v[k] = y1[k];
            call_slot(1);
            double x = it.v[0], y = it.v[1], z = it.v[2], w = it.v[3];
            double xx, yy;
            if (four) {
                xx = std::sqrt(x1[0] * x1[0] + x1[1] * x1[1] + x1[2] * x1[2] + x1[3] * x1[3]);
                yy = std::sqrt(x * x + y * y + z * z + w * w);
            } else {
                xx = std::sqrt(x1[0] * x1[0] + x1[1] * x1[1] + x1[2] * x1[2]);
                yy = std::sqrt(x * x + y * y + z * z);
            }
            xx = xx * s1 + yy * s2;
            double v[4] = {x1[0] * s1 + x * s2, x1[1] * s1 + y * s2, x1[2] * s1 + z * s2, x1[3] * s1 + w * s2};
            if (de && !four) v[3] = std::fabs(x1[3]) * s1 + std::fabs(w) * s2;
            double l2 = four ? v[0] * v[0] + v[1] * v[1] + v[2] * v[2] + v[3] * v[3]
                             : v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
            yy = xx / std::sqrt(l2 + 1e-40);
            v[0] *= yy;
            v[1] *= yy;
            v[2] *= yy;
            if (four) v[3] *= yy;
            for (int k = 0; k < 4; k++) it.v[k] = v[k];
            it.it_result += 1;
            it.rout = xx * xx;
            if (it.rout < it.otrap) it.otrap = it.rout;
            if (it.it_result >= it.max_it || it.rout > it.rstop) break;
        }
    }

    // ipol_start_3d (src/iteration.rs:403-415)
    void ipol_start_3d() {
        const double *c = it.c;
        start_j3();
        it.v[0] = c[0];
        it.v[1] = c[1];
        it.v[2] = c[2];
        it.rout = c[0] * c[0] + c[1] * c[1] + c[2] * c[2];
        it.otrap = it.rout;
    }

    // doInterpolHybridPas / Pas4D (src/iteration.rs:431-438,457-464)
    void ipol_plain() {
        if (it.four_d) {
            start_4d();
        } else {
            ipol_start_3d();
            it.v[3] = 0.0;
        }
        run_ipol(false);
        if (it.calc_sit) calc_smooth_iterations(0);
    }

    // doInterpolHybridPasDE, 3D (src/iteration.rs:441-455)
    double ipol_3d_de() {
        ipol_start_3d();
        it.v[3] = (it.de_option & 0x18) == 16 ? it.rout : 1.0;
        run_ipol(true);
        double w = it.v[3];
        double r = (it.de_option & 7) == 4 ? std::fabs(it.v[2]) * std::log(std::fabs(it.v[2])) / w
                                           : std::sqrt(it.rout) / std::fabs(w);
        if (it.calc_sit) calc_smooth_iterations(0);
        return r;
    }

    // mand_function (src/iteration.rs:859-866)
    void mand_function() {
        if (it.ipol)
            ipol_plain();
        else if (it.four_d)
            hybrid_4d();
        else
            hybrid_3d();
    }

    // Iteration::hybrid_3d (src/iteration.rs:521-537)
    void hybrid_3d() {
        const double *c = it.c;
        start_j3();
        it.v[0] = c[0];
        it.v[1] = c[1];
        it.v[2] = c[2];
        it.v[3] = 0.0;
        it.rout = c[0] * c[0] + c[1] * c[1] + c[2] * c[2];
        it.otrap = it.rout;
        int n = run();
        if (it.calc_sit) calc_smooth_iterations(n);
    }

    // Iteration::hybrid_3d_de (src/iteration.rs:539-585); de_option 32 is
    // the analytic power-8 extension with the textbook 0.5 r ln r / dr.
    double hybrid_3d_de() {
        const double *c = it.c;
        start_j3();
        it.v[0] = c[0];
        it.v[1] = c[1];
        it.v[2] = c[2];
        it.rout = c[0] * c[0] + c[1] * c[1] + c[2] * c[2];
        it.otrap = it.rout;
        switch (it.de_option & 0x38) {
            case 16: it.v[3] = it.rout; break;
            case 32: it.deriv1 = 1.0; break;
            default: it.v[3] = 1.0; break;
        }
        int n = run();
        double w = it.v[3];
        double result;
        if ((it.de_option & 0x38) == 32) {
            result = std::sqrt(it.rout) * 0.25 * std::log(it.rout) / it.deriv1;
        } else if ((it.de_option & 7) == 4) {
            result = std::fabs(it.v[1]) * std::log(std::fabs(it.v[1])) / w;
        } else {
            result = std::sqrt(it.rout) / std::fabs(w);
        }
        if (it.calc_sit) calc_smooth_iterations(n);
        return result;
    }

    // Marcher::calc_de_part for one part without dIFS (src/calc.rs:1495-1559)
    // Marcher::calc_de_part without dIFS (src/calc.rs:1495-1559) for the
    // current part; `invert`: the part as "inside" (kind 3's second part)
    double calc_de_part(bool invert) {
        double result;
        if (part.custom) {
            result = (it.ipol ? ipol_3d_de() : hybrid_3d_de()) * (double)part.dds;
        } else {
            mand_function();
            if (it.rout < 1e-200) {
                result = 0.0;
            } else {
                bool buf_sit = it.calc_sit;
                int buf_max_it = it.max_it;
                double buf_rout = it.rout;
                it.calc_sit = false;
                it.max_it = it.it_result;
                it.rstop = p.rstop3d;
                double off = (double)p.de_offset;
                double g[3];
                for (int k = 0; k < 3; k++) {
                    double bufd = it.c[k];
                    it.c[k] = bufd + off;
                    mand_function();
                    g[k] = (buf_rout - it.rout) * (buf_rout - it.rout);
                    it.c[k] = bufd;
                }
                result = buf_rout * std::log(buf_rout) * (double)part.dds /
                         (std::sqrt(g[0] + g[1] + g[2]) + (double)p.de_offset006);
                it.rout = buf_rout;
                it.it_result = it.max_it;
                it.max_it = buf_max_it;
                it.calc_sit = buf_sit;
                it.rstop = p.d_rstop;
            }
        }
        max_its_result = it.max_it;
        double mn = (double)ms_de_stop * 0.25;
        if (result < mn || std::isnan(result)) result = mn;
        if (invert) {
            result = (double)ms_de_stop * 4.0 - result * 3.0;
            if (result >= (double)ms_de_stop) it.it_result -= 1;
        }
        return result;
    }

    // set_part (src/calc.rs:1470-1490)
    void set_part(const PartO &pt) {
        part = pt;
        it.start_from = pt.start;
        it.end_to = pt.end;
        it.repeat_from = pt.repeat;
        it.max_it = pt.max_it;
        it.de_option = pt.de_option;
        it.four_d = pt.four;
    }

    // calc_de_fractal (src/calc.rs:1352-1456) without the mix (kind 6) and
    // inside rendering
    // Marcher::calc_de (src/calc.rs:1333-1455), ending in invert_inside
    double calc_de() {
        de_evals++;
        if (!p.decomb) return invert_inside(calc_de_part(false));
        double result = calc_de_part(false);
        const int b_it = it.it_result, b_max = max_its_result;
        const double b_otrap = it.otrap;
        const float b_sit = it.smooth_it;
        set_part(part2);
        double c0[3] = {it.c[0], it.c[1], it.c[2]};
        if (p.placed) {  // Place::apply: rotate_vector(p - offset, m)
            double d[3] = {c0[0] - p.place_offset[0], c0[1] - p.place_offset[1], c0[2] - p.place_offset[2]};
            for (int k = 0; k < 3; k++) it.c[k] = dot3(p.place_m[k], d);
        }
        double rst = calc_de_part(p.decomb == 3);
        if (p.placed) {
            rst *= p.place_scale;
            for (int k = 0; k < 3; k++) it.c[k] = c0[k];
        }
        const double ms = (double)ms_de_stop;
        bool restore1 = false;
        switch (p.decomb) {
            case 1:
                if (rst >= result)
                    restore1 = true;
                else
                    result = rst;
                break;
            case 2:
            case 3:
                if (rst > result)
                    result = rst;
                else
                    restore1 = true;
                break;
            default: {  // 4, 5: the smooth minima
                if (!(rst < result)) {
                    it.it_result = b_it;
                    max_its_result = b_max;
                }
                const double sm = p.decomb_smooth;
                if (p.decomb == 4) {
                    result = std::fmin(rst - std::fmax(sm - result, 0.0), result - std::fmax(sm - rst, 0.0));
                } else {
                    const double rd1 = std::fmax(sm - result - ms, 0.0);
                    const double d2 = std::fmax(sm - rst - ms, 0.0);
                    result = std::fmin(rst - rd1 * (sm - d2) / sm, result - d2 * (sm - rd1) / sm);
                }
                const double r2 = std::fabs(rst), w = std::fabs(result);
                const double f = 1.0 / (w + r2 + 1e-10);
                it.smooth_it = (float)(((double)it.smooth_it * w + (double)b_sit * r2) * f);
                it.otrap = (it.otrap * w + b_otrap * r2) * f;
                break;
            }
        }
        if (restore1) {
            it.it_result = b_it;
            it.smooth_it = b_sit;
            it.otrap = b_otrap;
            max_its_result = b_max;
        }
        set_part(part1);
        return invert_inside(result);
    }

    // invert_inside (src/calc.rs:1457-1472): inside rendering, the distance
    // to the surface from inside the set (no dIFS parts on this backend)
    double invert_inside(double r) {
        if (!calc_inside) return r;
        const double ms = (double)ms_de_stop;
        const double v = ms * 4.0 - r * 3.0;
        if (v >= ms) it.it_result -= 1;
        return v;
    }

    // calc_vgrads_fov (src/calc.rs:1561-1575), Common optic
    void calc_vgrads_fov(int ix) {
        double cafx = (double)((p.fovx_off - (float)ix) * p.fovx_mul);
        double v[3];
        build_view_vector_dfov(cafy, cafx, v);
        rot_rev(v, p.vgrads, vfov);
    }

    // calc_start_pos (src/calc.rs:1673-1690)
    void calc_start_pos(int ix, int iy) {
        double x = ix, y = iy;
        for (int k = 0; k < 3; k++) it.c[k] = p.ystart[k] + p.vgrads[0][k] * x + p.vgrads[1][k] * y;
    }

    void setup_ray(int x, int y) {
        cafy = ((double)y / (double)p.height - 0.5) * p.fov_y;
        calc_vgrads_fov(x + 1);
        calc_start_pos(x, y);
    }

    inline void step(double d) { add_weight(it.c, vfov, d); }

    // update_de_stop (src/calc.rs:1698-1703)
    inline void update_de_stop() {
        ms_de_stop = (float)((double)p.de_stop * (1.0 + mzz * (double)p.de_stop_factor));
    }

    // bin_search (src/calc.rs:1706-1735), without the experimental secant
    void bin_search(double &de, double last_step_width) {
        int itmp = p.de_add_steps;
        double dt1 = last_step_width * -0.5;
        while (std::fabs(de - (double)ms_de_stop) > 0.001) {
            mzz += dt1;
            step(dt1);
            update_de_stop();
            itmp -= 1;
            if (itmp <= 0) break;
            de = calc_de();
            if (it.it_result >= max_its_result) {
                dt1 = -std::fabs(dt1);
                continue;
            }
            double f = de - (double)ms_de_stop;
            dt1 = f < 0.0 ? std::fabs(dt1) * -0.55 : std::fabs(dt1) * 0.55;
        }
    }

    // bin_search_it (src/calc.rs:1737-1792); `steps` turns (< 0: the scene's)
    void bin_search_it(int steps = -1) {
        float yp = (float)it.max_it - 0.99f;
        int saved_max = it.max_it;
        it.max_it += 1;
        int itmp = steps < 0 ? p.de_add_steps : steps;
        it.calc_sit = true;
        double dt1 = 0.0, dmul = 1.0;
        bool first = true;
        float last_si = 0.0f, last_dif = 0.0f;
        for (;;) {
            mzz += dt1;
            step(dt1);
            calc_de();
            if (!first && last_dif < std::fabs(yp - it.smooth_it)) {
                mzz -= dt1;
                step(-dt1);
                it.smooth_it = last_si;
                if (dt1 > 0.0)
                    dmul *= 0.5;
                else
                    dmul *= 0.7;
            }
            last_dif = std::fabs(yp - it.smooth_it);
            last_si = it.smooth_it;
            if (it.smooth_it > (float)it.max_it - 0.1f) {
                dt1 = -3.0;
            } else {
                mzz -= 0.001;
                step(-0.001);
                calc_de();
                float r = last_si - it.smooth_it;
                if (std::fabs(r) < 1e-30f)
                    dt1 = (last_si < it.smooth_it ? 1.0 : 0.0) - 0.5;
                else if (r < 0.0f)
                    dt1 = (double)((yp - it.smooth_it) / (r * 500.0f));
                else
                    dt1 = (double)((yp - it.smooth_it) / (r * 1000.0f));
                if (dt1 > 4.0)
                    dt1 = std::sqrt(dt1) * 2.0;
                else if (dt1 < -9.0)
                    dt1 = std::sqrt(-dt1) * -3.0;
                dt1 = dt1 * dmul + 0.0005;
            }
            first = false;
            itmp -= 1;
            if (itmp < 0) break;
        }
        it.max_it = saved_max;
    }

    // calculate_normals with sm_normals = 0 (src/calc.rs:1813-1858); the
    // raw gradient goes to n, the return value is NN (smooth iterations)
    float calculate_normals(double n[3]) {
        const auto &vg = p.vgrads;
        it.calc_sit = true;
        double noffset = (double)std::fmin(p.de_stop, 1.0f) * (1.0 + mzz * (double)p.de_stop_factor) * 0.15;
        double ct1[3] = {it.c[0], it.c[1], it.c[2]};
        calc_de();
        float nn = it.smooth_it;
        nrm_it = it.it_result;
        it.calc_sit = false;
        const int order[3][2] = {{2, 2}, {0, 0}, {1, 1}};
        for (auto &ka : order) {
            int k = ka[0], axis = ka[1];
            std::memcpy(it.c, ct1, sizeof ct1);
            add_weight(it.c, vg[axis], noffset);
            double a = calc_de();
            add_weight(it.c, vg[axis], -2.0 * noffset);
            n[k] = (a - calc_de()) * 0.5;
        }
        std::memcpy(it.c, ct1, sizeof ct1);
        return nn;
    }

    // calc_roughness (src/calc.rs:1806-1811)
    static float calc_roughness(const double n[3], double dt2, double dsg) {
        double a = (dsg * 7.0 * dt2 * dt2 + 1e-40) / (n[0] * n[0] + n[1] * n[1] + n[2] * n[2] + 1e-40);
        return (float)std::fmin(std::fmax(std::sqrt(std::fmax(a, 0.0)) - 0.05, 0.0), 1.0);
    }

    // The smoothing of calculate_normals and calculate_normals_on_smooth_it
    // (src/calc.rs:1860-1901, 1958-1998): two sweeps of 2 sm samples across
    // the normal (create_xy_vecs_from_normals) around the centre value
    // `centre`, the roughness from their spread, and (sm < 8) the normal
    // moved by their mean slope.  `sample(sign)` evaluates at it.c and
    // returns the sweep's value: the DE's (value - centre), or the smooth
    // iteration's (centre - value).
    template <class Sample>
    void smooth_sweeps(double n[3], const double ct1[3], double noffset, Sample sample) {
        const auto &vg = p.vgrads;
        const int sm = p.sm_normals;
        double ssn = noffset * 3.0 / ((double)sm + 0.5);
        double dm = (double)sm * 2.0;
        double vx0[3], vy0[3], vx[3], vy[3];
        create_xy_vecs_from_normals(n, vx0, vy0);
        rot_rev(vx0, vg, vx);
        rot_rev(vy0, vg, vy);
        auto sweep = [&](const double dir[3], double &s1, double &ds) {
            std::memcpy(it.c, ct1, 3 * sizeof(double));
            add_weight(it.c, dir, -(double)sm * ssn);
            s1 = ds = 0.0;
            for (int k = -sm; k <= sm; k++) {
                if (k != 0) {
                    double d = sample() / (double)k;
                    s1 += d;
                    ds += d * d;
                }
                add_weight(it.c, dir, ssn);
            }
        };
        double nn1, ds1, nn2, ds2;
        sweep(vx, nn1, ds1);
        sweep(vy, nn2, ds2);
        double dsg = ds1 * dm - nn1 * nn1 + ds2 * dm - nn2 * nn2;
        double dt2 = noffset * 0.5 / (dm * ssn);
        s_roughness = calc_roughness(n, dt2, dsg);
        if (sm < 8) {
            n[0] += nn1 * dt2;
            n[1] += nn2 * dt2;
        }
    }

    // create_xy_vecs_from_normals (src/math.rs:218-232)
    static void create_xy_vecs_from_normals(const double n[3], double vx[3], double vy[3]) {
        double d = n[1] * n[1] + n[0] * n[0];
        if (d < 1e-50) {
            vx[0] = 1.0, vx[1] = 0.0, vx[2] = 0.0;
            vy[0] = 0.0, vy[1] = 1.0, vy[2] = 0.0;
            return;
        }
        double a = std::acos(std::fmin(std::fmax(-n[2] / std::sqrt(d + n[2] * n[2] + 1e-100), -1.0), 1.0)) * 0.5;
        double sin_a = std::sin(a), cos_a = std::cos(a);
        sin_a /= std::sqrt(d);
        double dd = -n[1] * sin_a, d1 = n[0] * sin_a;
        vx[0] = 1.0 - 2.0 * d1 * d1, vx[1] = 2.0 * dd * d1, vx[2] = 2.0 * d1 * cos_a;
        vy[0] = vx[1], vy[1] = 1.0 - 2.0 * dd * dd, vy[2] = -2.0 * dd * cos_a;
    }

    // calculate_normals for any smooth_normals (src/calc.rs:1813-1903): the
    // gradient (6 DEs, or the 124-point cube of smooth_normals 8), then the
    // smoothing sweeps and roughness; the raw normal to n, returns NN
    float calculate_normals_sm(double n[3]) {
        const auto &vg = p.vgrads;
        const int sm = p.sm_normals;
        if (sm == 0) return calculate_normals(n);
        it.calc_sit = true;
        double noffset = (double)std::fmin(p.de_stop, 1.0f) * (1.0 + mzz * (double)p.de_stop_factor) * 0.15;
        double ct1[3] = {it.c[0], it.c[1], it.c[2]};
        double dnn = calc_de();
        float nn = it.smooth_it;
        nrm_it = it.it_result;
        it.calc_sit = false;
        n[0] = n[1] = n[2] = 0.0;
        if (sm == 8) {
            double ssn = noffset * 1.3333;
            for (int a = -2; a <= 2; a++)
                for (int b = -2; b <= 2; b++)
                    for (int c = -2; c <= 2; c++) {
                        if ((a | b | c) == 0) continue;
                        std::memcpy(it.c, ct1, sizeof ct1);
                        add_weight(it.c, vg[2], (double)a * ssn);
                        add_weight(it.c, vg[1], (double)b * ssn);
                        add_weight(it.c, vg[0], (double)c * ssn);
                        double d = calc_de();
                        if (a != 0) n[2] += d / (double)a;
                        if (b != 0) n[1] += d / (double)b;
                        if (c != 0) n[0] += d / (double)c;
                    }
            for (int k = 0; k < 3; k++) n[k] *= 0.0075;
        } else {
            const int order[3][2] = {{2, 2}, {0, 0}, {1, 1}};
            for (auto &ka : order) {
                int k = ka[0], axis = ka[1];
                std::memcpy(it.c, ct1, sizeof ct1);
                add_weight(it.c, vg[axis], noffset);
                double a = calc_de();
                add_weight(it.c, vg[axis], -2.0 * noffset);
                n[k] = (a - calc_de()) * 0.5;
            }
        }
        noffset *= 2.0;
        if (sm < 8) {
            std::memcpy(it.c, ct1, sizeof ct1);
            add_weight(it.c, vg[0], -noffset);
            dnn += calc_de();
            add_weight(it.c, vg[0], 2.0 * noffset);
            dnn += calc_de();
            std::memcpy(it.c, ct1, sizeof ct1);
            add_weight(it.c, vg[1], -noffset);
            dnn += calc_de();
            add_weight(it.c, vg[1], 2.0 * noffset);
            dnn = (dnn + calc_de()) * 0.2;
        }
        smooth_sweeps(n, ct1, noffset, [&]() { return calc_de() - dnn; });
        std::memcpy(it.c, ct1, sizeof ct1);
        return nn;
    }

    // calculate_normals_on_smooth_it (src/calc.rs:1910-2003): the gradient
    // of the smooth iteration count (plain iterations, no DE), outside
    // rendering
    float calculate_normals_on_smooth_it(double n[3]) {
        const auto &vg = p.vgrads;
        const int sm = p.sm_normals;
        double noffset = (double)std::fmin(p.de_stop, 1.0f) * (1.0 + mzz * (double)p.de_stop_factor) * 0.15;
        it.calc_sit = true;
        double ct1[3] = {it.c[0], it.c[1], it.c[2]};
        mand_function();
        float nn = it.smooth_it;
        nrm_it = it.it_result;
        n[0] = n[1] = n[2] = 0.0;
        if (sm == 8) {
            double ssn = noffset * 1.3333;
            for (int a = -2; a <= 2; a++)
                for (int b = -2; b <= 2; b++)
                    for (int c = -2; c <= 2; c++) {
                        if ((a | b | c) == 0) continue;
                        std::memcpy(it.c, ct1, sizeof ct1);
                        add_weight(it.c, vg[2], (double)a * ssn);
                        add_weight(it.c, vg[1], (double)b * ssn);
                        add_weight(it.c, vg[0], (double)c * ssn);
                        mand_function();
                        double d = (double)it.smooth_it;
                        if (a != 0) n[2] -= d / (double)a;
                        if (b != 0) n[1] -= d / (double)b;
                        if (c != 0) n[0] -= d / (double)c;
                    }
            for (int k = 0; k < 3; k++) n[k] *= 0.0075;
        } else {
            const int order[3][2] = {{2, 2}, {0, 0}, {1, 1}};
            for (auto &ka : order) {
                int k = ka[0], axis = ka[1];
                std::memcpy(it.c, ct1, sizeof ct1);
                add_weight(it.c, vg[axis], noffset);
                mand_function();
                double a = (double)it.smooth_it;
                add_weight(it.c, vg[axis], -2.0 * noffset);
                mand_function();
                n[k] = ((double)it.smooth_it - a) * 0.5;
            }
        }
        if (sm > 0) {
            noffset *= 2.0;
            double acc = (double)nn;
            const int axes[4] = {0, 0, 1, 1};
            const double sgn[4] = {-1.0, 1.0, -1.0, 1.0};
            for (int q = 0; q < 4; q++) {
                std::memcpy(it.c, ct1, sizeof ct1);
                add_weight(it.c, vg[axes[q]], sgn[q] * noffset);
                mand_function();
                acc += (double)it.smooth_it;
            }
            nn = (float)(acc * 0.2);
            const double nnv = (double)nn;
            smooth_sweeps(n, ct1, noffset, [&]() {
                mand_function();
                return nnv - (double)it.smooth_it;
            });
        }
        std::memcpy(it.c, ct1, sizeof ct1);
        return nn;
    }

    // do_color_on_it (src/calc.rs:2010-2025): the colour from the start
    // vector (1) or from the orbit after color_on_it - 1 iterations
    void do_color_on_it() {
        if (p.color_on_it == 1) {
            it.v[0] = it.c[0];
            it.v[1] = it.c[1];
            it.v[2] = it.c[2];
        } else {
            int i = it.max_it;
            it.max_it = p.color_on_it - 1;
            it.rstop = p.rstop3d;
            calc_de();
            it.max_it = i;
            it.rstop = p.d_rstop;
        }
    }

    // store_normal (src/calc.rs:1794-1804)
    static void store_normal(int16_t out[3], const double n[3]) {
        double d = 32767.0 / std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2] + 1e-100);
        for (int k = 0; k < 3; k++) {
            double v = round_even(n[k] * d);
            out[k] = std::isnan(v) ? 0 : (int16_t)std::min(std::max(v, -32767.0), 32767.0);
        }
        if (out[0] == 0 && out[1] == 0) out[2] = out[2] > 0 ? 32767 : -32767;
    }

    // calc_zpos_and_rough (src/calc.rs:2194-2205): the depth code, and the
    // roughness in the low byte with smooth normals
    uint32_t zpos_fine() const {
        double zz = std::max(mzz, 0.0);
        double vr = round_even(p.zc_mul * (std::sqrt(zz * p.zcorr + 1.0) - 1.0));
        int64_t v = (int64_t)vr;
        int64_t itmp = std::min<int64_t>(std::max<int64_t>(8388352 - v, 0), 8388352);
        uint32_t r = (uint32_t)itmp << 8;
        if (p.sm_normals > 0) r |= (uint32_t)std::nearbyint(s_roughness * 255.0f) & 0xFFu;
        return r;
    }
};

// The G-buffer record (SiLight, src/gbuffer.rs:4-40) with the raw values
struct Px {
    int16_t normal[3] = {0, 0, 0};
    uint32_t zpos_fine = 32768u << 16;
    uint16_t shadow = 0, amb_shadow = 0, si_gradient = 0, otrap = 0;
    double depth = INFINITY;
    double n[3] = {0, 0, 0};
    uint16_t iters = 0;
    int32_t steps = 0;
};

void to_words(const Px &s, uint32_t w[5]) {
    w[0] = (uint32_t)(uint16_t)s.normal[0] | ((uint32_t)(uint16_t)s.normal[1] << 16);
    w[1] = (uint32_t)(uint16_t)s.normal[2] | ((uint32_t)s.shadow << 16);
    w[2] = s.zpos_fine;
    w[3] = (uint32_t)s.si_gradient | ((uint32_t)s.otrap << 16);
    w[4] = (uint32_t)s.amb_shadow;
}

void from_words(const uint32_t w[5], Px &s) {
    s.normal[0] = (int16_t)(uint16_t)(w[0] & 0xFFFF);
    s.normal[1] = (int16_t)(uint16_t)(w[0] >> 16);
    s.normal[2] = (int16_t)(uint16_t)(w[1] & 0xFFFF);
    s.shadow = (uint16_t)(w[1] >> 16);
    s.zpos_fine = w[2];
    s.si_gradient = (uint16_t)(w[3] & 0xFFFF);
    s.otrap = (uint16_t)(w[3] >> 16);
    s.amb_shadow = (uint16_t)(w[4] & 0xFFFF);
}

// do_color (src/calc.rs:2068-2090): the orbit trap (0), the last two radii
// (1), the angles of the orbit's end around the point (2-4), or its
// spherical angles (5, which also sets the gradient)
void do_color(const Marcher &m, Px &si) {
    const double x = m.it.v[0], y = m.it.v[1], z = m.it.v[2];
    const double *c = m.it.c;
    const double pi = 3.141592653589793;
    double s;
    switch (m.p.color_option) {
        case 1: s = std::log(m.it.rout / (m.it.rold + 1.0)) * (double)m.p.mct_color_mul; break;
        case 2: s = (std::atan2(y - c[1], x - c[0]) + pi) * 5200.0; break;
        case 3: s = (std::atan2(z - c[2], x - c[0]) + pi) * 5200.0; break;
        case 4: s = (std::atan2(z - c[2], y - c[1]) + pi) * 5200.0; break;
        case 5: {
            s = (std::atan2(x, y) + pi) * 5215.0;
            double r = std::sqrt(x * x + y * y + 1e-100 + z * z);
            double s2 = (pi + std::asin(std::fmin(std::fmax(z / r, -1.0), 1.0)) * 2.0) * 5215.0;
            si.si_gradient = clip15((float)s2);
            break;
        }
        default: s = m.it.otrap * 4096.0; break;
    }
    si.otrap = clip15((float)s);
}

// What do_color saw at the last surface the thread found (orc_color_probe):
// the orbit's end, the point, the radii and the two words.
struct ColorProbe {
    double v[3], c[3], rout, rold;
    uint16_t otrap, gradient;
    bool set;
};
thread_local ColorProbe *g_color_probe = nullptr;

// surface_found (src/calc.rs:2471-2517)
void surface_found(Marcher &m, Px &si, double dtmp, double last_step) {
    const mb3d_hip_params &p = m.p;
    bool de_limited = m.it.it_result < m.max_its_result || dtmp < (double)m.ms_de_stop;
    if (p.de_add_steps != 0) {
        if (de_limited)
            m.bin_search(dtmp, last_step);
        else
            m.bin_search_it();
    }
    si.iters = (uint16_t)std::max(m.it.it_result, 0);
    double n[3];
    float nn = p.normals_on_de ? m.calculate_normals_sm(n) : m.calculate_normals_on_smooth_it(n);
    Marcher::store_normal(si.normal, n);
    double l = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    for (int k = 0; k < 3; k++) si.n[k] = l > 0 ? n[k] / l : 0.0;
    float v;
    if (de_limited) {
        float lg = (float)std::log((double)std::fmax(p.de_stop, m.ms_de_stop) * p.step_width);
        v = 32767.0f - (nn + p.d_col_plus + p.col_var_de_stop_mul * lg) * p.mcts_m;
    } else {
        v = 32767.0f - nn * p.mcts_m;
    }
    si.si_gradient = clip15(v);
    if (p.color_on_it != 0) m.do_color_on_it();
    do_color(m, si);
    if (g_color_probe) {
        ColorProbe &o = *g_color_probe;
        for (int k = 0; k < 3; k++) {
            o.v[k] = m.it.v[k];
            o.c[k] = m.it.c[k];
        }
        o.rout = m.it.rout;
        o.rold = m.it.rold;
        o.otrap = si.otrap;
        o.gradient = si.si_gradient;
        o.set = true;
    }
    si.zpos_fine = m.zpos_fine();
    si.depth = m.mzz;
}

// march_from (src/calc.rs:2533-2657) without volumetric light and relax;
// the step cap is the declared extension (0: the reference's guard only)
void march_from(Marcher &m, Px &si, double dtmp, bool first_step, float step_count) {
    const mb3d_hip_params &p = m.p;
    float rsf_mul = 1.0f;
    double last_step = dtmp * (double)p.z_step_div;
    double last_de;
    uint32_t guard = 0;
    int32_t nsteps = 0;
    for (;;) {
        guard += 1;
        if (guard > 2000000) break;
        if (m.it.it_result >= m.max_its_result) {
            double dt1 = -0.5 * last_step;
            m.mzz += dt1;
            m.step(dt1);
            m.update_de_stop();
            dtmp = m.calc_de();
            last_step = -dt1;
        }
        if (m.it.it_result < p.min_it || (m.it.it_result < m.max_its_result && dtmp >= (double)m.ms_de_stop)) {
            if (p.step_cap > 0 && nsteps >= p.step_cap) break;  // extension: background
            last_de = dtmp;
            double d = std::fmax(0.11, (dtmp - (double)p.ms_de_sub * (double)m.ms_de_stop) * (double)p.z_step_div *
                                           (double)rsf_mul);
            double s1 = (double)(std::fmax(m.ms_de_stop, 0.4f) * p.mh04zsd);
            bool count = p.dfog_on_it == 0 || m.it.it_result == p.dfog_on_it;
            if (s1 < d) {
                if (count) step_count += (float)(s1 / d);
                d = s1;
            } else if (count) {
                step_count += 1.0f;
            }
            if (first_step) {
                first_step = false;
                m.seed = (int32_t)((uint32_t)m.seed * 214013u + 2531011u);
                d *= (double)(m.seed & 0x7FFFFFFF) * (1.0 / (double)0x7FFFFFFF);
            }
            m.mzz += d;
            if (m.mzz > p.zend) break;  // background
            nsteps++;
            last_step = d;
            m.step(d);
            m.update_de_stop();
            dtmp = m.calc_de();
            if (dtmp > last_de + last_step) dtmp = last_de + last_step;
            if (last_de > dtmp + 1e-30) {
                double t = last_step / (last_de - dtmp);
                rsf_mul = t < 1.0 ? (float)std::fmax(0.5, t) : 1.0f;
            } else {
                rsf_mul = 1.0f;
            }
        } else {
            surface_found(m, si, dtmp, last_step);
            break;
        }
    }
    float sc = std::fmin(std::fmax(step_count, 0.0f), 1023.0f);
    si.shadow = (uint16_t)std::nearbyint(sc);
    si.steps = nsteps;
}

// The cone march (experimental in the reference: src/cone.rs:1-31,
// src/calc.rs:2357-2435) with one level of blocks, bw x bh pixels: the
// block's coarse ray from the start plane.
constexpr double kConeMinStep = 0.05, kConeMinShare = 0.1, kConeMargin = 1.5;  // src/calc.rs:2659-2674

struct ConeCell {
    double t = 0.0;  // the safe depth (step widths); +inf: the block is background
    float count = 0.0f;  // the dynamic fog's step count up to there
};

// Marcher::ray (src/calc.rs:2328-2355), Common optic: the ray of the
// picture's point (fx, fy), fractions allowed
void frac_ray(const mb3d_hip_params &p, double fx, double fy, double c[3], double v[3]) {
    double cafy = (fy / (double)p.height - 0.5) * p.fov_y;
    double cafx = (double)((p.fovx_off - (float)(fx + 1.0)) * p.fovx_mul);
    double w[3];
    build_view_vector_dfov(cafy, cafx, w);
    rot_rev(w, p.vgrads, v);
    for (int k = 0; k < 3; k++) c[k] = p.ystart[k] + p.vgrads[0][k] * fx + p.vgrads[1][k] * fy;
}

inline double dist3(const double a[3], const double b[3]) {
    double d[3] = {a[0] - b[0], a[1] - b[1], a[2] - b[2]};
    return std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
}

// cone_march (src/calc.rs:2357-2435) of the block at (bx bw, by bh), t0 = 0
ConeCell cone_march(Marcher &m, int bx, int by, int bw, int bh) {
    const mb3d_hip_params &p = m.p;
    double x0 = (double)(bx * bw), y0 = (double)(by * bh);
    double c0[3], v0[3];
    frac_ray(p, x0 + ((double)bw - 1.0) * 0.5, y0 + ((double)bh - 1.0) * 0.5, c0, v0);
    double a = 0.0, slope = 0.0;
    for (int j = 0; j < bh; j++)
        for (int i = 0; i < bw; i++) {
            double c[3], v[3];
            frac_ray(p, x0 + (double)i, y0 + (double)j, c, v);
            a = std::fmax(a, dist3(c, c0));
            slope = std::fmax(slope, dist3(v, v0));
        }
    a /= p.step_width;
    slope /= p.step_width;
    for (int k = 0; k < 3; k++) {
        m.it.c[k] = c0[k];
        m.vfov[k] = v0[k];
    }
    m.mzz = 0.0;
    m.it.calc_sit = false;
    m.max_its_result = p.max_it;
    double lip = kConeMargin / (double)std::fmin(std::fmax(p.z_step_div, 0.05f), 1.0f);
    double div = 1.0 + slope * lip + (double)p.de_stop * (double)p.de_stop_factor;
    double last_de = INFINITY, last_step = 0.0, t = 0.0;
    ConeCell out;
    float count = 0.0f;
    for (int g = 0; g < 2000000; g++) {
        m.update_de_stop();
        double de = m.calc_de();
        if (m.it.it_result >= m.max_its_result) break;
        if (de > last_de + last_step) de = last_de + last_step;
        double rsf = last_step > 0.0 && last_de > de + 1e-30 ? std::fmin(std::fmax(last_step / (last_de - de), 0.5), 1.0)
                                                            : 1.0;
        double stop = (double)m.ms_de_stop;
        double reach = (a + slope * t) * lip + stop;
        double cap = (double)(std::fmax(m.ms_de_stop, 0.4f) * p.mh04zsd);
        double sd = std::fmin((de - reach - (double)p.ms_de_sub * stop) * (double)p.z_step_div * rsf / div, cap);
        double d = std::fmax(0.11, (de - (double)p.ms_de_sub * stop) * (double)p.z_step_div * rsf);
        if (sd < kConeMinStep * stop || sd < kConeMinShare * std::fmin(d, cap)) break;
        if (p.dfog_on_it == 0 || m.it.it_result == p.dfog_on_it) count += (float)(std::fmin(sd, p.zend - t) / d);
        if (t + sd > p.zend) {
            t = INFINITY;
            break;
        }
        t += sd;
        m.mzz = t;
        m.step(sd);
        last_de = de;
        last_step = sd;
    }
    out.t = t;
    out.count = count;
    return out;
}

// march_pixel / march_pixel_from without cut planes, inside rendering or
// ground (src/calc.rs:2207-2322); with a cone cell (`cone`), mode 1: its
// background verdict, mode 2: also the march from its depth (a landing in
// the surface marches the pixel from the start)
Px march_pixel(Marcher &m, int x, int y, const ConeCell *cone = nullptr, int mode = 0) {
    const mb3d_hip_params &p = m.p;
    Px si;
    si.amb_shadow = 5000;
    m.setup_ray(x, y);
    m.it.calc_sit = false;
    m.mzz = 0.0;
    m.ms_de_stop = p.de_stop;
    bool first_step = p.first_step_random != 0;
    double dtmp = m.calc_de();
    if (m.it.it_result >= m.max_its_result || dtmp < (double)m.ms_de_stop) {
        // inside the set at the start plane: calc_nangles_for_cut(0)
        si.zpos_fine = 0x7FFF0000u;
        si.normal[0] = 0;
        si.normal[1] = 0;
        si.normal[2] = -32767;
        si.n[2] = -1.0;
        do_color(m, si);
        double t = std::fmin(std::fmax(m.it.rout / p.d_rstop, 0.0), 1.0);
        si.si_gradient = (uint16_t)(32768 + (int)round_even(32767.0 * t));
        si.depth = 0.0;
        si.iters = (uint16_t)std::max(m.it.it_result, 0);
        return si;
    }
    if (cone && cone->t > 0.0) {
        if (cone->t > p.zend) {  // background, with the count of the way (src/calc.rs:2299-2305)
            si.shadow = (uint16_t)std::nearbyint(std::fmin(std::fmax(cone->count, 0.0f), 1023.0f));
            return si;
        }
        if (mode == 2) {
            double d = cone->t;
            m.mzz = d;
            m.step(d);
            m.update_de_stop();
            dtmp = m.calc_de();
            if (dtmp < (double)m.ms_de_stop || m.it.it_result >= m.max_its_result) return march_pixel(m, x, y);
            march_from(m, si, dtmp, first_step, cone->count);
            return si;
        }
    }
    march_from(m, si, dtmp, first_step, 0.0f);
    return si;
}

// surface depth of a G-buffer pixel, the middle of its depth code's last
// digit (src/calc.rs:2167, 2748-2750)
inline double depth_of_code(const mb3d_hip_params &p, double zf, double frac) {
    double a = (frac - zf) / p.zc_mul + 1.0;
    return (a * a - 1.0) / p.zcorr;
}

// hs_surface_start (src/calc.rs:2741-2765)
bool hs_surface_start(Marcher &m, int x, int y, const Px &si) {
    const mb3d_hip_params &p = m.p;
    if ((si.zpos_fine >> 16) >= 32768 || si.si_gradient >= 32768) return false;
    m.setup_ray(x, y);
    double zf = (double)(si.zpos_fine >> 8);
    m.mzz = depth_of_code(p, zf, 8388351.5);
    add_weight(m.it.c, m.vfov, m.mzz);
    m.update_de_stop();
    double d = m.calc_de();
    bool de_limited = m.it.it_result < m.max_its_result || d < (double)m.ms_de_stop;
    if (de_limited) {
        double dt1 = depth_of_code(p, zf, 8388351.9) - m.mzz;
        m.bin_search(d, dt1);
    } else {
        m.bin_search_it();
    }
    return true;
}

// dot_normalize (src/calc.rs:2734-2737)
inline double dot_normalize(const double a[3], const double b[3]) {
    return dot3(a, b) / std::fmax(std::sqrt(dot3(a, a) * dot3(b, b)), 1e-300);
}

// hard_shadow_row for one pixel, directional lights, without the row
// neighbour shortcut (src/calc.rs:2777-2899; the card's variant)
void hard_shadow_pixel(Marcher &m, int x, int y, Px &si, const mb3d_hip_post_job &job) {
    const mb3d_hip_params &p = m.p;
    uint16_t clear_mask = 0;
    for (int l = 0; l < job.n_lights; l++) clear_mask |= (uint16_t)(0x400 << job.light_idx[l]);
    si.shadow &= (uint16_t)~clear_mask;
    double nv[3] = {si.normal[0] / 32767.0, si.normal[1] / 32767.0, si.normal[2] / 32767.0};
    double nvec[3];
    rot_rev(nv, p.vgrads, nvec);
    if (!hs_surface_start(m, x, y, si)) return;
    m.mzz = std::fmax(m.mzz - 0.1, 0.0);
    add_weight(m.it.c, m.vfov, -0.1);
    m.update_de_stop();
    double ic[3] = {m.it.c[0], m.it.c[1], m.it.c[2]};
    // hs_max_len (src/calc.rs:2767-2775)
    float mz = (float)std::fmin(m.mzz, p.zend * 0.4);
    double max_lhs = (double)((float)(p.width + y) * 0.6f *
                              (1.0f + 0.5f * mz * (float)std::fmax(p.fov_y, 0.0) / (float)p.height) *
                              job.shadow_max_len_mul);
    for (int l = 0; l < job.n_lights; l++) {
        int bit = 0x400 << job.light_idx[l];
        std::memcpy(m.it.c, ic, sizeof ic);
        double zz2 = m.mzz;
        const double *lvec = job.light_vec[l];
        double zz2mul = -dot_normalize(lvec, m.vfov);
        if (dot3(nvec, lvec) > 0.0) {
            si.shadow |= (uint16_t)bit;
            continue;
        }
        double dt1 = max_lhs;
        if (dt1 > 0.0) {
            m.ms_de_stop = (float)((double)p.de_stop * (1.0 + std::fabs(zz2) * (double)p.de_stop_factor));
            float rsf = 2.0f;
            double d = m.calc_de();
            int guard = 0;
            for (;;) {
                guard += 1;
                if (guard > 1000000) break;
                double last_de = d;
                float a = std::fmax(0.11f, (float)((d - (double)p.ms_de_sub * (double)m.ms_de_stop) *
                                                   (double)p.z_step_div * (double)rsf));
                float b = std::fmax(m.ms_de_stop, 0.4f) * p.mh04zsd;
                double st = (double)std::fmin(a, b);
                dt1 -= st;
                add_weight(m.it.c, lvec, -st);
                zz2 += st * zz2mul;
                m.ms_de_stop = (float)((double)p.de_stop * (1.0 + std::fabs(zz2) * (double)p.de_stop_factor));
                d = m.calc_de();
                if (m.it.it_result >= m.max_its_result || d <= (double)m.ms_de_stop) break;
                if (d > last_de + st) d = last_de + st;
                if (last_de > d + 1e-30) {
                    float t = (float)(st / (last_de - d));
                    rsf = t < 1.0f ? std::fmax(t, 0.5f) : 1.0f;
                } else {
                    rsf = 1.0f;
                }
                if (dt1 < 0.0) break;
            }
        }
        if (dt1 > 0.0) si.shadow |= (uint16_t)bit;
    }
}

// ---------------------------------------------------------------------------
// DE ambient occlusion (src/deao.rs)

typedef float SVec[3];

// ray_dir (src/deao.rs:35-40)
void ray_dir(double ya, double za, float out[3]) {
    double sy = std::sin(ya), cy = std::cos(ya);
    double sz = std::sin(za), cz = std::cos(za);
    out[0] = (float)(sy * cz);
    out[1] = (float)(-sy * sz);
    out[2] = (float)(-cy);
}

// normal_matrix (src/deao.rs:42-58)
void normal_matrix(const float n[3], float m[3][3]) {
    double a = std::acos(std::fmin(std::fmax(-(double)n[2], -1.0), 1.0)) * 0.5;
    double sa = std::sin(a), ca = std::cos(a);
    double nn = std::sqrt((double)(n[1] * n[1] + n[0] * n[0]));
    double q[4];
    if (nn < 1e-25) {
        q[0] = q[1] = q[2] = 0.0;
        q[3] = ca;
    } else {
        sa /= nn;
        q[0] = -(double)n[1] * sa;
        q[1] = (double)n[0] * sa;
        q[2] = 0.0;
        q[3] = ca;
    }
    double r[3][3] = {
        {1.0 - 2.0 * (q[1] * q[1] + q[2] * q[2]), 2.0 * (q[0] * q[1] + q[2] * q[3]), 2.0 * (q[0] * q[2] - q[1] * q[3])},
        {2.0 * (q[0] * q[1] - q[2] * q[3]), 1.0 - 2.0 * (q[0] * q[0] + q[2] * q[2]), 2.0 * (q[2] * q[1] + q[0] * q[3])},
        {2.0 * (q[0] * q[2] + q[1] * q[3]), 2.0 * (q[1] * q[2] - q[0] * q[3]), 1.0 - 2.0 * (q[0] * q[0] + q[1] * q[1])},
    };
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) m[i][j] = (float)r[i][j];
}

inline float sdot(const float a[3], const float b[3]) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

struct Rays {
    std::vector<std::array<float, 3>> dirs;
    int quality = 0;
    int row_count[4] = {0, 0, 0, 0};
    float d_step_mul, d_min_a_dif, correction_weight;
};

// build_rays (src/deao.rs:75-105); quality 4 is the 8-ray set (extension):
// the quality-1 geometry with 7 rays on its ring
Rays build_rays(int quality) {
    Rays r;
    const double TAU = 6.283185307179586;
    int q = std::min(quality, 4);
    r.quality = q;
    if (q == 0) {
        double abr = 60.0 * (3.141592653589793 / 180.0);
        for (int x = 0; x < 3; x++) {
            std::array<float, 3> d;
            ray_dir(0.5 * abr, x * TAU / 3.0, d.data());
            r.dirs.push_back(d);
        }
        r.d_step_mul = 1.8f;
        r.d_min_a_dif = -1.0f;
        r.correction_weight = 0.3f;
        return r;
    }
    int qg = q == 4 ? 1 : q;
    std::array<float, 3> d0;
    ray_dir(0.0, 0.0, d0.data());
    r.dirs.push_back(d0);
    double abr = 1.5707963267948966 / ((double)qg + 0.9);
    for (int y = 1; y <= qg; y++) {
        double dt1 = y * abr;
        int n = (int)std::nearbyint(std::sin(dt1) * TAU / abr);
        if (q == 4) n = 7;
        r.row_count[y] = n;
        for (int x = 0; x < n; x++) {
            std::array<float, 3> d;
            ray_dir(dt1, x * TAU / n, d.data());
            r.dirs.push_back(d);
        }
    }
    r.d_step_mul = (float)(1.0 + std::sin(abr));
    r.d_min_a_dif = (float)std::cos(abr * 1.2);
    r.correction_weight = qg == 1 ? 0.2f : 0.1666f;
    return r;
}

// The ray directions of pixel (x, y) with dithering 1 (2x2) or 2 (3x3):
// the rings shifted by the pixel's place in its cell, and the centre ray
// left out on the shifted rows (src/deao.rs:318-342)
std::vector<std::array<float, 3>> dither_rays(const Rays &base, int dither, int x, int y) {
    const double TAU = 6.283185307179586;
    const double dt1 = (double)(y % (dither + 1)) * 0.5 / (double)dither;
    const double dt2 = (double)(x % (dither + 1)) * 0.5 / (double)dither;
    std::vector<std::array<float, 3>> dirs;
    std::array<float, 3> d;
    if (base.quality == 0) {
        for (int i = 0; i < 3; i++) {
            ray_dir((dt1 + 0.5) * (50.0 * (3.141592653589793 / 180.0)), ((double)i + dt2) * TAU / 3.0, d.data());
            dirs.push_back(d);
        }
        return dirs;
    }
    const double abr = 1.5707963267948966 / ((double)base.quality + 0.9);
    if (!(dt1 > 0.1)) {
        ray_dir(0.0, 0.0, d.data());
        dirs.push_back(d);
    }
    for (int r = 1; r <= base.quality; r++) {
        const int n = base.row_count[r];
        for (int i = 0; i < n; i++) {
            ray_dir(abr * ((double)r + dt1 - 0.25), ((double)i + dt2) * TAU / (double)n, d.data());
            dirs.push_back(d);
        }
    }
    return dirs;
}

// The row's random numbers for the first AO steps: DeaoRows::row's seed and
// get_rand (src/deao.rs:262-270), a 32-bit linear congruential generator
int32_t deao_row_seed(int y) { return (int32_t)(uint32_t)(0x24563487LL + ((long long)y + 1) * 0x324594A1LL); }

double deao_rand(int32_t &s) {
    s = (int32_t)((uint32_t)s * 0x343FDu + 0x269EC3u);
    return (double)(((uint32_t)s >> 8) & 0x7FFFFFu) / (double)0x7FFFFF;
}

// surface_point (src/calc.rs:2157-2192)
bool surface_point(Marcher &m, int x, int y, const Px &si, int steps, double thr) {
    const mb3d_hip_params &p = m.p;
    if ((si.zpos_fine >> 16) >= 32768 || si.si_gradient >= 32768) return false;
    m.setup_ray(x, y);
    m.it.calc_sit = false;
    double zf = (double)(si.zpos_fine >> 8);
    m.mzz = depth_of_code(p, zf, 8388351.5);
    add_weight(m.it.c, m.vfov, m.mzz);
    m.update_de_stop();
    double d = std::fmin(m.calc_de(), (double)m.ms_de_stop * 2.0);
    if (m.it.it_result < m.max_its_result || d < (double)m.ms_de_stop) {
        double dt1 = depth_of_code(p, zf, 8388351.9) - m.mzz;
        int itmp = steps;
        while (itmp > 0 && std::fabs(d - (double)m.ms_de_stop) > thr) {
            m.mzz += dt1;
            add_weight(m.it.c, m.vfov, dt1);
            m.update_de_stop();
            itmp -= 1;
            if (itmp > 0) {
                d = m.calc_de();
                dt1 = d < (double)m.ms_de_stop ? std::fabs(dt1) * -0.55 : std::fabs(dt1) * 0.55;
            }
        }
    } else {
        m.bin_search_it();
        m.ms_de_stop = (float)((double)p.de_stop * (1.0 + std::fabs(m.mzz) * (double)p.de_stop_factor));
    }
    return true;
}

// What deao_pixel computed for one pixel, for the tests that check it
// against brute force (orc_deao_probe).
struct DeaoProbe {
    double ic[3];        // the rebuilt surface point
    double dirs[33][3];  // each ray's step in world units (step_ao-scaled directions)
    double min_ra[33];   // each ray's openness (times de_mul)
    double first[33];    // each ray's first step over step_ao * d_step_mul (1 without the random first step)
    double step_ao, max_dist, ms_ao, md_d10, de_mul, dsm;
    int count;
    // the walk of one ray (trace_ray >= 0): its samples' distances and DEs
    int trace_ray = -1, trace_cap = 0, trace_n = 0;
    double *trace_t = nullptr, *trace_de = nullptr;
    float ms_walk = 0.0f;  // the marcher's DE stop during the walks (the DE's floor is a quarter of it)
};

// deao_row for one pixel (src/deao.rs:268-418); the marcher's parameters
// have de_add_steps = 5.  `rnd`: the row's generator state, which the random
// first steps advance once per ray of every object pixel, left to right
void deao_pixel(Marcher &m, int x, int y, Px &si, const Rays &base, const mb3d_hip_post_job &job,
                int32_t *rnd = nullptr, DeaoProbe *probe = nullptr) {
    const mb3d_hip_params &p = m.p;
    double sw = p.step_width;
    double rot_v[3][3];
    normalise_matrix_to(sw, p.vgrads, rot_v);
    double s_max_d = (double)job.deao_max_len * 0.5 *
                     std::sqrt((double)p.height * (double)p.height + (double)p.width * (double)p.width);
    double de_stop = (double)p.de_stop;
    si.amb_shadow = 0;
    if ((si.zpos_fine >> 16) >= 32768 || si.si_gradient >= 32768) return;
    if (!surface_point(m, x, y, si, 5, 0.004)) return;
    double ms_de_stop = std::fmin((double)m.ms_de_stop, 1e6);
    double ic[3] = {m.it.c[0], m.it.c[1], m.it.c[2]};
    double step_ao = ms_de_stop / de_stop;
    double max_dist = s_max_d * std::sqrt(step_ao);
    double dsm = (double)base.d_step_mul;
    if (p.de_stop_factor == 0.0f)
        ms_de_stop = de_stop / (dsm * dsm);
    else
        ms_de_stop = ms_de_stop / (dsm * dsm);
    const std::vector<std::array<float, 3>> dirs =
        job.deao_dither > 0 ? dither_rays(base, job.deao_dither, x, y) : base.dirs;
    int ray_count = (int)dirs.size();
    double de_mul = std::sqrt((double)ray_count * 0.5);
    double abr = 1.2 / std::asin(std::fmin(std::fmax(1.0 / de_mul, -1.0), 1.0));
    double md_d10 = 0.1 / (max_dist * de_mul);
    float nvf[3] = {(float)si.normal[0], (float)si.normal[1], (float)si.normal[2]};
    float l = std::fmax(std::sqrt(sdot(nvf, nvf)), 1e-30f);
    float n[3] = {nvf[0] / l, nvf[1] / l, nvf[2] / l};
    float rot_w[3][3];
    normal_matrix(n, rot_w);
    std::vector<float> min_ra(ray_count, 0.0f);
    for (int k = 0; k < ray_count; k++) {
        const float *d = dirs[k].data();
        float v[3] = {sdot(rot_w[0], d), sdot(rot_w[1], d), sdot(rot_w[2], d)};
        double vd[3] = {(double)v[0], (double)v[1], (double)v[2]};
        double sv[3];
        rot_rev(vd, rot_v, sv);
        double dt1 = step_ao * dsm;
        double s_tmp = 1.0;
        bool b_end = false;
        bool b_first = job.deao_first_step_random != 0;
        if (probe) probe->first[k] = 1.0;
        for (;;) {
            if (b_first) {
                b_first = false;
                const double f = deao_rand(*rnd) * 1.5 + 0.5;
                dt1 *= f;
                if (probe) probe->first[k] = f;
            } else if (dt1 > max_dist) {
                dt1 = max_dist;
                b_end = true;
            }
            std::memcpy(m.it.c, ic, sizeof ic);
            add_weight(m.it.c, sv, dt1);
            double dt2 = m.calc_de();
            if (probe && k == probe->trace_ray && probe->trace_n < probe->trace_cap) {
                probe->trace_t[probe->trace_n] = dt1;
                probe->trace_de[probe->trace_n++] = dt2;
            }
            s_tmp = std::fmin(s_tmp, (dt2 - ms_de_stop + dt1 * md_d10) / dt1);
            if (s_tmp < 0.02) break;
            dt1 += std::fmax(dt2, dt1 * dsm);
            if (b_end) break;
        }
        min_ra[k] = (float)(std::fmax(s_tmp, 0.0) * de_mul);
        if (probe) {
            for (int i = 0; i < 3; i++) probe->dirs[k][i] = sv[i];
            probe->min_ra[k] = min_ra[k];
        }
    }
    if (probe) {
        for (int i = 0; i < 3; i++) probe->ic[i] = ic[i];
        probe->step_ao = step_ao;
        probe->max_dist = max_dist;
        probe->ms_ao = ms_de_stop;
        probe->md_d10 = md_d10;
        probe->de_mul = de_mul;
        probe->dsm = dsm;
        probe->count = ray_count;
        probe->ms_walk = m.ms_de_stop;
    }
    double amount = 0.0;
    for (int k = 0; k < ray_count; k++) {
        float s_add = 0.0f;
        if (min_ra[k] < 1.0f) {
            float max_add = 1.0f - min_ra[k];
            for (int k2 = 0; k2 < ray_count; k2++) {
                if (k2 == k) continue;
                float dt = sdot(dirs[k].data(), dirs[k2].data());
                if (dt > base.d_min_a_dif) {
                    float ac = std::acos(std::fmin(std::fmax(dt, -1.0f), 1.0f));
                    float overlap = min_ra[k2] - (float)((double)ac * abr) + 1.0f;
                    if (overlap > 0.0f) s_add += std::fmin(max_add, overlap) * base.correction_weight;
                }
            }
        }
        amount += (double)std::fmin(s_add + min_ra[k], 1.0f);
    }
    double a = round_even(16383.0 * (1.0 - amount / (double)ray_count));
    si.amb_shadow = (uint16_t)std::fmax(a, 0.0);
}

int nthreads(int t) {
    if (t > 0) return t;
    unsigned h = std::thread::hardware_concurrency();
    return h ? (int)h : 1;
}

template <class F>
void for_rows_interleaved(int rows, int threads, F f) {
    threads = std::max(1, std::min(threads, rows));
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; t++)
        pool.emplace_back([=]() {
            for (int r = t; r < rows; r += threads) f(r);
        });
    for (auto &th : pool) th.join();
}

int check_params(const mb3d_hip_params *p) {
    if (!p) return fail("params is NULL");
    if (p->abi_version != MB3D_HIP_ABI_VERSION) return fail("params ABI version mismatch");
    if (p->width <= 0 || p->height <= 0) return fail("bad picture size");
    if (p->n_slots < 1 || p->n_slots > MB3D_HIP_MAX_SLOTS) return fail("bad slot count");
    return 0;
}

// The DE AO's dithering: 0..2, not with the 8-ray set
int check_job(const mb3d_hip_post_job &job) {
    if (job.deao && (job.deao_dither < 0 || job.deao_dither > 2)) return fail("bad DE AO dithering");
    if (job.deao && job.deao_dither && job.deao_quality >= 4) return fail("no dithering of the 8-ray set");
    return 0;
}

void run_post(const mb3d_hip_params &p, const mb3d_hip_post_job &job, int row0, int rows, std::vector<Px> &px,
              int threads) {
    int w = p.width;
    if (job.n_lights > 0) {
        for_rows_interleaved(rows, nthreads(threads), [&](int r) {
            Marcher m(p);
            int y = row0 + r;
            for (int x = 0; x < w; x++) hard_shadow_pixel(m, x, y, px[(size_t)r * w + x], job);
        });
    }
    if (job.deao) {
        mb3d_hip_params pc = p;
        pc.de_add_steps = 5;  // DeaoRows::new (src/deao.rs:254-260)
        Rays base = build_rays(job.deao_quality);
        for_rows_interleaved(rows, nthreads(threads), [&](int r) {
            Marcher m(pc);
            int y = row0 + r;
            int32_t rnd = deao_row_seed(y);
            for (int x = 0; x < w; x++) deao_pixel(m, x, y, px[(size_t)r * w + x], base, job, &rnd);
        });
    }
}

}  // namespace

// ---------------------------------------------------------------------------
// parameter set-up: CalcParams::new (src/calc.rs:375-898) for the built-in
// formulas, one alternating 3D hybrid, no DE combination or interpolation

namespace {

int formula_de_option(const mb3d_hip_formula &f) {  // Formula::de_option (src/formulas.rs:492-499)
    if (f.kind == MB3D_HIP_FORMULA_QUATERNION) return 4;
    return f.kind == MB3D_HIP_FORMULA_AMAZINGBOX ? 11 : 0;
}
double formula_de_scale(const mb3d_hip_formula &f) {  // de_scale (src/formulas.rs:502-509)
    if (f.kind == MB3D_HIP_FORMULA_AMAZINGBOX || f.kind == MB3D_HIP_FORMULA_BULBOX) return 0.2;
    return f.kind == MB3D_HIP_FORMULA_FOLDINGINTPOW ? 0.5 : 1.0;
}
double formula_default_rstop(const mb3d_hip_formula &f) {  // default_rstop (src/formulas.rs:520-528)
    const int k = f.kind;
    return (k == MB3D_HIP_FORMULA_AMAZINGBOX || k == MB3D_HIP_FORMULA_BULBOX || k == MB3D_HIP_FORMULA_FOLDINGINTPOW)
               ? 1024.0
               : 16.0;
}
double formula_si_pow(const mb3d_hip_formula &f) {  // si_pow (src/formulas.rs:531-545)
    if (f.kind == MB3D_HIP_FORMULA_INTPOW) return (double)std::min(std::max(f.power, 2), 8);
    if (f.kind == MB3D_HIP_FORMULA_REALPOWER || f.kind == MB3D_HIP_FORMULA_AEXIONC)
        return f.options[0] == 0.0 ? 1e-40 : f.options[0];
    return 2.0;
}
double formula_first_option(const mb3d_hip_formula &f) {  // first_option (src/formulas.rs:547-558)
    switch (f.kind) {
        case MB3D_HIP_FORMULA_INTPOW: return f.z_mul;
        case MB3D_HIP_FORMULA_AMAZINGBOX: return f.scale;
        case MB3D_HIP_FORMULA_FOLDINGINTPOW: return f.options[1];  // the Z multiplier
        default: return f.options[0];
    }
}
bool formula_first_option_double(const mb3d_hip_formula &f) {  // first_option_type (src/formulas.rs:560-568)
    return f.kind != MB3D_HIP_FORMULA_INTPOW && f.kind != MB3D_HIP_FORMULA_FOLDINGINTPOW;
}

// Rust's f64::round (half away from zero) as set_option rounds integer
// options (src/formulas.rs:383-393)
int round_clamp(double v, int lo, int hi) {
    double r = std::round(v);
    if (!(r >= (double)lo)) return lo;  // (NaN: Rust's cast gives 0, which the clamp raises to lo)
    if (r > (double)hi) return hi;
    return (int)r;
}

// The options of the kinds other than Integer Power and Amazing Box as the
// formula holds them: Folding Int Pow's power an integer 2..8, Aexion C's
// switches 0/1 and its mode an integer 0..31 (set_option,
// src/formulas.rs:377-404).  `power` gets Folding Int Pow's power.
void normalized_options(const mb3d_hip_formula &f, double o[8], int32_t &power) {
    for (int k = 0; k < MB3D_HIP_FORMULA_OPTIONS; k++) o[k] = f.options[k];
    if (f.kind == MB3D_HIP_FORMULA_FOLDINGINTPOW) {
        power = round_clamp(f.options[0], 2, 8);
        o[0] = (double)power;
    } else if (f.kind == MB3D_HIP_FORMULA_AEXIONC) {
        for (int k : {2, 3, 6}) o[k] = std::round(f.options[k]) != 0.0 ? 1.0 : 0.0;
        o[7] = (double)round_clamp(f.options[7], 0, 31);
    }
}

// build_smatrix4 (src/math.rs:178-204): XW, then YW, then ZW plane
// rotations, kept in f32
void build_smatrix4(double xw, double yw, double zw, double r[4][4]) {
    auto ini = [](double m[4][4]) {
        for (int i = 0; i < 4; i++)
            for (int k = 0; k < 4; k++) m[i][k] = i == k ? 1.0 : 0.0;
    };
    auto mul = [](const double a[4][4], const double b[4][4], double o[4][4]) {
        double t[4][4];
        for (int i = 0; i < 4; i++)
            for (int k = 0; k < 4; k++) {
                double acc = 0.0;  // Iterator::sum adds in order (the sign of an all-zero sum aside)
                for (int q = 0; q < 4; q++) acc += a[i][q] * b[q][k];
                t[i][k] = acc;
            }
        std::memcpy(o, t, sizeof t);
    };
    double sx = std::sin(xw), cx = std::cos(xw);
    double sy = std::sin(yw), cy = std::cos(yw);
    double sz = std::sin(zw), cz = std::cos(zw);
    double m[4][4], m2[4][4], m3[4][4];
    ini(m);
    m[0][0] = cx;
    m[3][3] = cx;
    m[0][3] = -sx;
    m[3][0] = sx;
    ini(m2);
    m2[1][1] = cy;
    m2[3][3] = cy;
    m2[1][3] = -sy;
    m2[3][1] = sy;
    mul(m2, m, m2);
    ini(m3);
    m3[2][2] = cz;
    m3[3][3] = cz;
    m3[2][3] = -sz;
    m3[3][2] = sz;
    mul(m3, m2, r);
    for (int i = 0; i < 4; i++)
        for (int k = 0; k < 4; k++) r[i][k] = (double)(float)r[i][k];
}

}  // namespace

extern "C" int32_t orc_params_from_scene(const mb3d_hip_scene *sc, mb3d_hip_params *out) {
    if (!sc || !out) return fail("NULL argument");
    if (sc->abi_version != MB3D_HIP_ABI_VERSION) return fail("scene ABI version mismatch");
    int nf = sc->n_formulas;
    if (nf < 1 || nf > MB3D_HIP_MAX_SLOTS) return fail("1..6 formulas");
    for (int i = 0; i < nf; i++) {
        int k = sc->formulas[i].kind;
        if (k < MB3D_HIP_FORMULA_INTPOW || k > MB3D_HIP_FORMULA_AEXIONC) return fail("unknown formula kind");
    }
    // interpolation hybrid: the first two formulas, weights normalised to a
    // sum of 1 in f32 (src/calc.rs:378-397)
    const bool ipol = sc->interpolation != 0;
    float ipw[2] = {1.0f, 0.0f};
    if (ipol) {
        if (nf < 2) return fail("an interpolation hybrid needs two formulas");
        float x1 = sc->interpolation_w[0], x2 = sc->interpolation_w[1];
        float y1 = x1 + x2;
        if (y1 < 1e-10f) {
            y1 = 1.0f;
            x1 = 1.0f;
            x2 = 0.0f;
        } else {
            y1 = 1.0f / y1;
        }
        ipw[0] = x1 * y1;
        ipw[1] = x2 * y1;
    }
    bool any = false;
    for (int i = 0; i < nf; i++) any = any || sc->formulas[i].iterations != 0;
    if (!any && !ipol) return fail("no formula with iterations > 0");
    mb3d_hip_params p;
    std::memset(&p, 0, sizeof p);
    p.abi_version = MB3D_HIP_ABI_VERSION;
    int width = sc->width, height = sc->height;
    p.width = width;
    p.height = height;
    // slots; the last with iterations ends the hybrid (CheckHybridOptions)
    int last = ipol ? 1 : nf - 1;  // an interpolation hybrid's two slots count one iteration each
    while (!ipol && last > 0 && sc->formulas[last].iterations == 0) last--;
    int ns = last + 1;
    p.n_slots = ns;
    for (int i = 0; i < ns; i++) {
        const mb3d_hip_formula &f = sc->formulas[i];
        p.slot_kind[i] = f.kind;
        p.slot_iters[i] = ipol ? 1 : std::abs(f.iterations);
        p.slot_uncounted[i] = !ipol && f.iterations < 0;
        p.slot_power[i] = f.power;
        p.slot_z_mul[i] = f.z_mul;
        p.slot_scale[i] = f.scale;
        p.slot_min_r[i] = f.min_r;
        p.slot_fold[i] = f.fold;
        normalized_options(f, p.slot_opt[i], p.slot_power[i]);
    }
    // the DE combination's two parts (src/calc.rs:421-431): only when both
    // have slots with iterations
    const int dck = ipol ? 0 : sc->decomb;
    const bool dc = dck != 0 && sc->decomb_end1 < last && sc->decomb_start2 <= last;
    if (dc && (dck < 1 || dck > 6)) return fail("unknown DE combination");
    int end_to = dc ? std::min(sc->decomb_end1, last) : last;
    int repeat_from = std::min(sc->repeat_from, end_to);
    while (repeat_from > 0 && p.slot_iters[repeat_from] <= 0) repeat_from--;
    // combine_de_options (src/calc.rs:318-372)
    auto combine = [&](int start, int end) {
        int de_option = -1;
        for (int i = start; i <= end; i++) {
            if (p.slot_iters[i] == 0) continue;
            int de = formula_de_option(sc->formulas[i]);
            if (de < 0 || de == de_option) continue;
            if (de_option == -1 || de_option == 21 || de_option == 22)
                de_option = de;
            else if (de_option == 2)
                de_option = (de == 2 || de == 5 || de == 6 || de == 11) ? 2 : 0;
            else if (de_option == 4)
                de_option = (de == 5 || de == 6) ? 4 : 0;
            else if (de_option == 5)
                de_option = de == 4 ? 4 : ((de == 2 || de == 11) ? 2 : (de != 6 ? 0 : 5));
            else if (de_option == 6)
                de_option = (de == 2 || de == 4 || de == 5 || de == 11) ? de : 0;
            else if (de_option == 11)
                de_option = (de == 2 || de == 5) ? 2 : (de != 6 ? 0 : 11);
            // (other options keep theirs)
        }
        return de_option;
    };
    int de_option = combine(0, end_to);
    // the second part (src/calc.rs:487-520)
    int start2 = 0, end2 = 0, rep2 = 0, de2 = -1;
    if (dc) {
        start2 = std::min(sc->decomb_start2, last);
        end2 = std::min(std::max(sc->decomb_end2, start2), last);
        de2 = combine(start2, end2);
        if (de2 < 0) return fail("DE combination: the second formula part has no valid DE option");
        if (dck == 6 && de2 != 20) return fail("DE combination 'mix': the second hybrid part must be dIFS");
        rep2 = std::min(std::max(sc->decomb_repeat2, start2), end2);
        while (rep2 > start2 && p.slot_iters[rep2] <= 0) rep2--;
    }
    const bool is_custom_de2 = de2 == 2 || de2 == 5 || de2 == 6 || de2 == 11 || de2 == 20;
    if (ipol) {
        // ipol_de_option (src/calc.rs:255-307): the formulas whose raw weight
        // is positive (as the bits of the f32)
        de_option = -1;
        for (int x = 0; x < 2; x++) {
            int32_t bits;
            std::memcpy(&bits, &sc->interpolation_w[x], 4);
            if (bits <= 0) continue;
            int de = formula_de_option(sc->formulas[x]);
            if (de < 0 || de > 20 || de == de_option) continue;
            if ((de_option >= 0 && de_option < 20 && de > 19) || (de_option > 19 && de >= 0 && de < 20)) {
                de_option = -1;
                break;
            }
            if (de_option == -1)
                de_option = de;
            else if (de_option == 2)
                de_option = (de == 2 || de == 5 || de == 6 || de == 11) ? 2 : 0;
            else if (de_option == 4)
                de_option = (de == 5 || de == 6) ? 4 : 0;
            else if (de_option == 5)
                de_option = de == 4 ? 4 : ((de == 2 || de == 11) ? 2 : (de != 6 ? 0 : 5));
            else if (de_option == 6)
                de_option = (de == 2 || de == 4 || de == 5 || de == 11) ? de : 0;
            else if (de_option == 11)
                de_option = (de == 2 || de == 5) ? 2 : (de != 6 ? 0 : 11);
        }
        if (de_option > 19) de_option = 20;
        if (de_option >= 20) return fail("dIFS formulas do not work in an interpolation hybrid");
        if (de_option < 0) return fail("formula option is not valid");
    }
    if (sc->de_mode == MB3D_HIP_DE_ANALYTIC_P8) {
        if (ipol) return fail("analytic p8 DE needs one Integer Power slot");
        if (ns != 1 || sc->formulas[0].kind != MB3D_HIP_FORMULA_INTPOW) return fail("analytic p8 DE needs one Integer Power slot");
        de_option = 32;
    } else if (sc->de_mode != MB3D_HIP_DE_REFERENCE) {
        return fail("unknown DE mode");
    }
    bool is_custom_de = de_option == 2 || de_option == 5 || de_option == 6 || de_option == 11 || de_option == 20 ||
                        de_option == 32;
    p.de_option = de_option;
    p.is_custom_de = is_custom_de;
    p.four_d = de_option >= 4 && de_option <= 6;  // HybridMode::Alt4D / Ipol4D (src/calc.rs:475-484)
    p.ipol = ipol;
    p.ipol_w[0] = ipw[0];
    p.ipol_w[1] = ipw[1];
    p.do_julia = sc->julia != 0;
    for (int k = 0; k < 4; k++) p.julia_c[k] = sc->julia_c[k];
    build_smatrix4(sc->rot_4d[0], sc->rot_4d[1], sc->rot_4d[2], p.smatrix4);
    if (dc) {
        p.decomb = dck;
        p.start2 = start2;
        p.end2 = end2;
        p.repeat2 = rep2;
        p.max_it2 = sc->decomb_iterations2;
        p.de_option2 = de2;
        p.is_custom_de2 = is_custom_de2;
        p.four_d2 = de2 >= 4 && de2 <= 6;
    }
    for (int i = 0; i < ns; i++) {
        p.slot_ade[i] = dc && i > end_to ? is_custom_de2 : is_custom_de;
        int de = formula_de_option(sc->formulas[i]);
        p.slot_uncounted[i] = p.slot_uncounted[i] || de < 0 || de > 20;
    }
    // DE scale, colour helpers, smooth-it logs
    int max_it = sc->iterations;
    float fhln_default = (float)(1.0 / std::log(2.0));
    for (int k = 0; k < 6; k++) p.fhln[k] = fhln_default;
    double n = 0, x1 = 0, y1 = 0, ii = 0;
    bool have_part1 = false;
    double part1_scale = 1.0;  // the first part's, once the second part's slots begin
    for (int x = 0; x < ns; x++) {
        if (p.slot_iters[x] == 0) continue;
        const mb3d_hip_formula &f = sc->formulas[x];
        const bool second = dc && x > end_to;
        if (second && !have_part1) {
            have_part1 = true;
            part1_scale = n > 0.0 && de_option != 20 ? x1 / n : 1.0;
            n = x1 = y1 = ii = 0.0;
        }
        const bool custom = second ? is_custom_de2 : is_custom_de;
        if (ipol || !p.slot_uncounted[x]) {
            double sip = formula_si_pow(f);
            double z1 = sip > 1.0 ? sip : formula_first_option(f);
            p.fhln[x] = (float)(1.0 / std::log(std::fmax(2.0, std::fabs(z1))));
        }
        // interpolation hybrid: weighted by the weights, counted once
        double z1 = ipol ? (double)ipw[x] : (double)p.slot_iters[x], cnt = ipol ? 1.0 : (double)p.slot_iters[x];
        int j = formula_de_option(f);
        if (custom && (ipol || !p.slot_uncounted[x]) && (j == 5 || j == 11)) {
            n += cnt;
            ii += cnt;
            double x2 = formula_first_option(f);
            if (x2 == 0.0) x2 = 1e-40;
            if (x2 < 0.0)
                x1 += 0.65 * z1;
            else
                x1 += (x2 * 1.2 + 1.0) / x2 * z1;
            y1 += std::fabs(x2) * z1;
        } else if (j >= 0 || ipol) {
            n += cnt;
            x1 += (custom ? 1.0 : formula_de_scale(f)) * z1;  // ade_scale is 1 for built-ins
        }
    }
    float d_col_plus = 0.0f;
    double d_de_scale = 1.0, d_de_scale2 = 1.0;
    if (have_part1) {
        d_de_scale = part1_scale;
        if (n > 0.0 && de2 != 20) d_de_scale2 = x1 / n;
    } else if (n > 0.0 && de_option != 20) {
        d_de_scale = ipol ? x1 : x1 / n;
    }
    if (n > 0.0 && ii > 0.0) d_col_plus = (float)(std::pow(std::fabs(y1 / ii), 0.25) * 40.0 - 49.0);
    // mctColVarDEstopMul (src/calc.rs:594-650); with a DE combination the
    // value of the part with more counted iterations
    float col_var = 0.6f;
    {
        double yy1 = 0, zz1 = 0, zz2 = 0;
        int ia = 0, itmp = 0;
        bool second = false;
        for (int x = 0; x < ns; x++) {
            bool counted = ipol || (p.slot_iters[x] > 0 && !p.slot_uncounted[x]);
            if (!counted) {
                if (p.slot_iters[x] > 0) {
                    if (x > end_to)
                        ia++;
                    else
                        itmp++;
                }
                continue;
            }
            if (x > end_to && !second) {
                second = true;
                col_var = zz1 > 0.0 ? (float)(yy1 / zz1) : 0.6f;
                zz2 = zz1;
                zz1 = 0.0;
                yy1 = 0.0;
                if (!dc || de2 == 20) break;
            }
            const mb3d_hip_formula &f = sc->formulas[x];
            double x2 = ipol ? (double)ipw[x] : (double)p.slot_iters[x];
            int de = formula_de_option(f);
            if (de == 0 || de == 4) {
                yy1 += x2 * 0.6;
                zz1 += x2;
            } else if ((de >= 1 && de <= 3) || (de >= 5 && de <= 40)) {
                if (formula_first_option_double(f)) {
                    double v = formula_first_option(f);
                    if (v > 1.0) {
                        if (v < 1.2) v = 1.1 + (v - 1.0) * (v - 1.0) * 2.5;
                        yy1 += x2 * 0.5 / (std::log(v) + 0.03);
                        zz1 += x2;
                    } else if (v < -1.0) {
                        if (v > -1.2) v = -1.1 - (v + 1.0) * (v + 1.0) * 2.5;
                        yy1 += x2 * 0.5 / (std::log(-v) + 0.03);
                        zz1 += x2;
                    }
                }
            }
        }
        if ((zz1 >= zz2 && zz1 > 1e-3) || (ia > itmp && std::fabs(zz1 - zz2) < 1e-3 && zz1 > 1e-3))
            col_var = (float)(yy1 / zz1);
        else if (zz2 < 1e-3)
            col_var = 0.6f;
    }
    // effective_rstop (src/scene.rs:799-809)
    double rstop = sc->rstop;
    if (!(rstop > 0.0)) {
        rstop = 0.0;
        for (int i = 0; i < nf; i++)
            if (sc->formulas[i].iterations > 0) rstop = std::fmax(rstop, formula_default_rstop(sc->formulas[i]));
        rstop = std::fmax(rstop, 1.0001);
    }
    double step_width = 2.1345 / (sc->zoom * (double)width);  // Scene::step_width (src/scene.rs:717-719)
    double fov_y = sc->fov_y * (3.141592653589793 / 180.0);
    float z_step_div0 = (float)std::fmax(sc->z_step_div, 0.0001);
    float de_stop = (float)std::fmax(sc->de_stop, 0.001);
    double x1c = std::fmin(std::fmax(fov_y * 0.5, 0.01), 1.5);
    p.pl_optic_z = (float)(std::cos(x1c) * x1c / std::sin(x1c));
    p.fovx_off = 0.5f * (float)width;  // stereo_xoff() = 0.5 for a mono picture
    p.fovx_mul = (float)(fov_y / (double)height);
    // CalcPPZvals
    double zcorr = std::sin(std::fmax(sc->fov_y * (3.141592653589793 / 180.0), 1.0) / (double)height);
    double zc_mul =
        32767.0 * 256.0 / (std::sqrt((sc->z_end - sc->z_start) * zcorr / step_width + 1.0) - 0.999999999);
    double zend = std::fmax(1e-10, (sc->z_end - sc->z_start) / step_width);
    float mh04zsd = (float)((double)std::max(width, height) * 0.5 * std::sqrt((double)z_step_div0 + 0.001) *
                            std::fmax(0.01, sc->raystep_limiter));
    // GetDEstopFactor
    float de_stop_factor = 0.0f;
    if (sc->vary_de_stop_on_fov) {
        double ze = std::fmax(1e-16, (sc->z_end - sc->z_start) / step_width);
        double xx1 = 0.01 * (double)height /
                     (std::sin(0.01) * std::fmax(1.0 / 65535.0, sc->fov_y * (3.141592653589793 / 180.0)));
        double xx2 = step_width * (xx1 + ze) / xx1;
        de_stop_factor = (float)((xx2 - step_width) / (step_width * std::fmax(1e-6, ze)));
    }
    double xh = (double)width * 0.5, yh = (double)height * 0.5;
    double z1 = (sc->z_start - sc->mid[2]) / step_width;
    normalise_matrix_to(step_width, sc->vgrads, p.vgrads);
    for (int k = 0; k < 3; k++)
        p.ystart[k] = sc->mid[k] + z1 * p.vgrads[2][k] - yh * p.vgrads[1][k] - xh * p.vgrads[0][k];
    float de_offset = std::fmin(de_stop * 0.1f, 0.004f);
    float dds;
    if (is_custom_de)
        dds = (float)((double)(float)d_de_scale / step_width);
    else
        dds = (float)d_de_scale * de_offset;
    if (dc) {
        p.d_de_scale2 = is_custom_de2 ? (float)((double)(float)d_de_scale2 / step_width) : (float)d_de_scale2 * de_offset;
        p.decomb_smooth = (double)std::fmax(sc->decomb_smooth, 1e-30f) / step_width;
        // the second part as an object of its own (src/calc.rs:748-752)
        const bool moved = sc->part2_offset[0] != 0.0 || sc->part2_offset[1] != 0.0 || sc->part2_offset[2] != 0.0 ||
                           sc->part2_rotation[0] != 0.0 || sc->part2_rotation[1] != 0.0 ||
                           sc->part2_rotation[2] != 0.0 || sc->part2_scale != 1.0;
        if (sc->part2_placed && moved && dck != 6) {
            const double d2r = 3.141592653589793 / 180.0;  // f64::to_radians
            const double xa = sc->part2_rotation[0] * d2r, ya = sc->part2_rotation[1] * d2r,
                         za = sc->part2_rotation[2] * d2r;
            // build_rot_matrix (src/math.rs:466-475)
            const double sx = std::sin(xa), cx = std::cos(xa), sy = std::sin(ya), cy = std::cos(ya);
            const double sz = std::sin(za), cz = std::cos(za);
            const double rm[3][3] = {{cy * cz, -cy * sz, sy},
                                     {sx * sy * cz + cx * sz, cx * cz - sx * sy * sz, -sx * cy},
                                     {sx * sz - cx * sy * cz, cx * sy * sz + sx * cz, cx * cy}};
            normalise_matrix_to(1.0 / std::fmax(sc->part2_scale, 1e-300), rm, p.place_m);
            for (int k = 0; k < 3; k++) p.place_offset[k] = sc->part2_offset[k];
            p.place_scale = sc->part2_scale;
            p.placed = 1;
        }
    }
    float de_offset006 = de_offset * 0.06f;
    de_offset = (float)((double)de_offset * step_width);
    float z_step_div = z_step_div0, ms_de_sub = 0.0f;
    if (sc->step_sub_de_stop) {
        float s = z_step_div0;
        s = s * s + (1.2f * s) * (1.0f - s);
        z_step_div = s;
        ms_de_sub = std::fmin(std::sqrt(s), 0.9f);
    }
    float ln_rstop = (float)std::log(std::log(rstop));
    double d_rstop = (double)(float)(rstop * rstop);
    double rstop3d = (double)(float)((d_rstop * d_rstop) * 64.0);
    // the iterations of both parts of a DE combination (iMaxitF2)
    int its_all = max_it;
    if (dc) its_all = dck > 5 ? max_it + sc->decomb_iterations2 : std::max(max_it, sc->decomb_iterations2);
    float mcts_m = 32767.0f / (float)std::max(its_all + 1, 1);
    d_col_plus = d_col_plus + (float)its_all * 0.1f;

    p.d_de_scale = dds;
    p.sm_normals = std::min(std::max(sc->smooth_normals, 0), 8);
    p.de_stop = de_stop;
    p.z_step_div = z_step_div;
    p.de_add_steps = sc->bin_search_steps;
    p.fov_y = fov_y;
    p.max_it = max_it;
    p.min_it = std::min(sc->min_iterations, max_it);
    p.color_option = std::min(sc->color_option, 5);
    p.color_on_it = std::min(std::max(sc->color_on_it, 0), 255);  // (a u8 in CalcParams, src/calc.rs:27)
    p.mct_color_mul = (float)(sc->color_mul * 512.0);
    p.d_rstop = d_rstop;
    p.rstop3d = rstop3d;
    p.dfog_on_it = sc->dfog_on_it;
    p.d_col_plus = d_col_plus;
    p.zcorr = zcorr;
    p.zc_mul = zc_mul;
    p.zend = zend;
    p.step_width = step_width;
    p.col_var_de_stop_mul = col_var;
    p.first_step_random = sc->first_step_random;
    p.mh04zsd = mh04zsd;
    p.de_stop_factor = de_stop_factor;
    p.ln_rstop = ln_rstop;
    p.normals_on_de = sc->normals_on_de || is_custom_de;
    p.de_offset = de_offset;
    p.de_offset006 = de_offset006;
    p.mcts_m = mcts_m;
    p.ms_de_sub = ms_de_sub;
    p.end_to = end_to;
    p.repeat_from = repeat_from;
    p.step_cap = std::max(sc->step_cap, 0);
    *out = p;
    return 0;
}

extern "C" const char *orc_last_error(void) { return g_err.c_str(); }

// pixel_seed (src/calc.rs:1125-1140)
extern "C" int32_t orc_pixel_seed(int64_t y, int64_t x) {
    uint32_t row = (uint32_t)(int32_t)(int64_t)(0x24563487LL + (y + 1) * 0x324594A1LL);
    uint32_t a = 214013u, c = 2531011u, ra = 1u, rc = 0u;
    uint64_t n = x > 0 ? (uint64_t)x : 0;
    while (n > 0) {
        if (n & 1) {
            rc = rc * a + c;
            ra = ra * a;
        }
        c = c * a + c;
        a = a * a;
        n >>= 1;
    }
    return (int32_t)(ra * row + rc);
}

extern "C" void orc_iterate(const mb3d_hip_formula *f, double v[4], double j[4], double rout, int32_t ade) {
    double o[8];
    int32_t power = f->power;
    normalized_options(*f, o, power);
    switch (f->kind) {
        case MB3D_HIP_FORMULA_INTPOW: int_pow(f->power, f->z_mul, v, j); break;
        case MB3D_HIP_FORMULA_AMAZINGBOX: amazing_box(f->scale, f->min_r, f->fold, v, j, ade != 0); break;
        case MB3D_HIP_FORMULA_REALPOWER: real_power(o[0], o[1], v, j, rout); break;
        case MB3D_HIP_FORMULA_QUATERNION: quaternion(o[0], o[1], v, j); break;
        case MB3D_HIP_FORMULA_TRICORN: tricorn(o[0], o[1], v, j); break;
        case MB3D_HIP_FORMULA_BULBOX: bulbox(o, v, j, rout); break;
        case MB3D_HIP_FORMULA_FOLDINGINTPOW:
            for (int k = 0; k < 3; k++) v[k] = box_fold(v[k], o[2]);
            int_pow(power, o[1], v, j);
            break;
        case MB3D_HIP_FORMULA_AEXIONC: aexion_c(o, v, j); break;
        default: break;
    }
}

extern "C" double orc_de_at(const mb3d_hip_params *p, const double pos[3], int32_t *its) {
    if (check_params(p)) return NAN;
    Marcher m(*p);
    m.it.c[0] = pos[0];
    m.it.c[1] = pos[1];
    m.it.c[2] = pos[2];
    double d = m.calc_de();
    if (its) *its = m.it.it_result;
    return d * p->step_width;
}

extern "C" int32_t orc_de_many(const mb3d_hip_params *p, const double *pos, int64_t n, double *de, int32_t *its,
                               int32_t threads) {
    if (int e = check_params(p)) return e;
    const int64_t chunk = 4096;
    const int64_t nchunks = (n + chunk - 1) / chunk;
    for_rows_interleaved((int)nchunks, nthreads(threads), [&](int c) {
        Marcher m(*p);
        for (int64_t i = (int64_t)c * chunk; i < std::min(n, (int64_t)(c + 1) * chunk); i++) {
            for (int k = 0; k < 3; k++) m.it.c[k] = pos[3 * i + k];
            de[i] = m.calc_de();
            its[i] = m.it.it_result;
        }
    });
    return 0;
}

extern "C" void orc_pixel_ray(const mb3d_hip_params *p, int32_t x, int32_t y, double c[3], double v[3]) {
    Marcher m(*p);
    m.setup_ray(x, y);
    for (int k = 0; k < 3; k++) {
        c[k] = m.it.c[k];
        v[k] = m.vfov[k];
    }
}

// One pixel's DE AO (deao_pixel) on its card record `rec` (5 words), with
// what it computed: out = the surface point (3), the ray steps in world
// units (33 x 3), the openness per ray (33), then step_ao, max_dist, the
// DE stop of the rays, md_d10, de_mul and d_step_mul, then each ray's first
// step factor (33); returns the ray count (0: no surface), the amb word in
// *amb.  `rnd`: the row generator's state before the pixel (advanced), for
// the random first step.
extern "C" int32_t orc_deao_probe(const mb3d_hip_params *pp, const mb3d_hip_post_job *job, int32_t x, int32_t y,
                                  const uint32_t rec[5], double *out, int32_t *amb, int32_t *rnd) {
    if (check_params(pp) || !job || !job->deao) return -1;
    if (job->deao_first_step_random && !rnd) return -1;
    if (check_job(*job)) return -1;
    mb3d_hip_params pc = *pp;
    pc.de_add_steps = 5;  // DeaoRows::new (src/deao.rs:254-260)
    Rays base = build_rays(job->deao_quality);
    Marcher m(pc);
    Px si;
    from_words(rec, si);
    DeaoProbe pr{};
    deao_pixel(m, x, y, si, base, *job, rnd, &pr);
    *amb = si.amb_shadow;
    for (int i = 0; i < 3; i++) out[i] = pr.ic[i];
    for (int k = 0; k < 33; k++)
        for (int i = 0; i < 3; i++) out[3 + 3 * k + i] = pr.dirs[k][i];
    for (int k = 0; k < 33; k++) out[102 + k] = pr.min_ra[k];
    const double info[6] = {pr.step_ao, pr.max_dist, pr.ms_ao, pr.md_d10, pr.de_mul, pr.dsm};
    for (int i = 0; i < 6; i++) out[135 + i] = info[i];
    for (int k = 0; k < 33; k++) out[141 + k] = pr.first[k];
    return pr.count;
}

// The walk of ray `ray` of one pixel's DE AO (deao_pixel on the card record
// `rec`): the samples' distances along the ray in steps (t) and the DE the
// walk took at each (de), at most `cap`; out2 = the marcher's DE stop during
// the walk (the DE's floor is a quarter of it) and the ray's openness;
// returns the sample count (-1: bad arguments or no such ray).
extern "C" int32_t orc_deao_walk(const mb3d_hip_params *pp, const mb3d_hip_post_job *job, int32_t x, int32_t y,
                                 const uint32_t rec[5], int32_t *rnd, int32_t ray, int32_t cap, double *t,
                                 double *de, double out2[2]) {
    if (check_params(pp) || !job || !job->deao || ray < 0 || cap <= 0 || !t || !de) return -1;
    if (job->deao_first_step_random && !rnd) return -1;
    if (check_job(*job)) return -1;
    mb3d_hip_params pc = *pp;
    pc.de_add_steps = 5;  // DeaoRows::new (src/deao.rs:254-260)
    Rays base = build_rays(job->deao_quality);
    Marcher m(pc);
    Px si;
    from_words(rec, si);
    DeaoProbe pr{};
    pr.trace_ray = ray;
    pr.trace_cap = cap;
    pr.trace_t = t;
    pr.trace_de = de;
    deao_pixel(m, x, y, si, base, *job, rnd, &pr);
    if (ray >= pr.count) return -1;
    out2[0] = (double)pr.ms_walk;
    out2[1] = pr.min_ra[ray] / pr.de_mul;
    return pr.trace_n;
}

// The DE AO's row generator: the seed of row y, then n states and numbers
extern "C" int32_t orc_deao_row_seed(int32_t y) { return deao_row_seed(y); }

extern "C" void orc_deao_rand(int32_t seed, int32_t n, int32_t *states, double *vals) {
    for (int i = 0; i < n; i++) {
        vals[i] = deao_rand(seed);
        states[i] = seed;
    }
}

// create_xy_vecs_from_normals (src/math.rs:218-232), for the tests
extern "C" void orc_xy_vecs(const double n[3], double vx[3], double vy[3]) {
    Marcher::create_xy_vecs_from_normals(n, vx, vy);
}

extern "C" double orc_march_ray(const mb3d_hip_params *p, const double origin[3], const double dir[3], double zz,
                                int32_t *its, int32_t *steps) {
    if (check_params(p)) return NAN;
    Marcher m(*p);
    for (int k = 0; k < 3; k++) {
        m.it.c[k] = origin[k];
        m.vfov[k] = dir[k];
    }
    m.mzz = zz;
    m.it.calc_sit = false;
    m.update_de_stop();
    double d = m.calc_de();
    Px si;
    march_from(m, si, d, false, 0.0f);
    if (its) *its = si.iters;
    if (steps) *steps = si.steps;
    return si.depth;
}

// SrTracer::march_ray (src/reflect.rs:1464-1674) for one task, with the
// pass's options at the card's: no volumetric light, no ground plane, no
// hard shadows or occlusion at the hit (amb_shadow 5000), outside
// rendering.  The view ray of the task's pixel (setup_ray) gives zzplus.
static void march_ray_task(Marcher &m, const mb3d_hip_ray_task &t, bool depth_sqrt, mb3d_hip_ray_hit &h) {
    const mb3d_hip_params &p = m.p;
    const double sw = p.step_width;
    const int px = (int)(t.pixel % (uint32_t)p.width), py = (int)(t.pixel / (uint32_t)p.width);
    m.setup_ray(px, py);  // m.vfov: the pixel's view vector
    const double new_vec[3] = {(double)t.dir[0], (double)t.dir[1], (double)t.dir[2]};
    const double max_l = t.max_len > 0.0f ? (double)t.max_len : p.zend;
    double zz = (double)t.zz;
    for (int k = 0; k < 3; k++) m.it.c[k] = t.origin[k];
    m.it.calc_sit = false;
    m.it.max_it = p.max_it;
    const double zzplus = dot3(m.vfov, new_vec) / std::sqrt(dot3(m.vfov, m.vfov) * dot3(new_vec, new_vec) + 1e-100);
    const float last_zz = (float)std::fabs(zz);
    double zz2 = std::fmin((double)p.de_stop * 0.25, 1.0);
    zz += zz2 * zzplus;
    add_weight(m.it.c, new_vec, zz2);
    auto upd = [&](double z) { m.ms_de_stop = (float)((double)p.de_stop * (1.0 + std::fabs(z) * (double)p.de_stop_factor)); };
    upd(zz);
    bool open_air = true, de_limited = true;
    float step_count = 0.0f;
    double rlast_sw = 0.0;
    if (t.inside) {
        zz2 = p.zend;
    } else {
        open_air = false;
        float rsfmul = 1.0f;
        double dtmp = m.calc_de();
        rlast_sw = dtmp * (double)p.z_step_div;
        uint32_t guard = 0;
        if (dtmp > 1e-10) {
            for (;;) {
                const double rlast_de = dtmp;
                double dstep = std::fmax((dtmp - (double)p.ms_de_sub * (double)m.ms_de_stop) * (double)p.z_step_div * (double)rsfmul,
                                         0.11);
                const double s = (double)(std::fmax(m.ms_de_stop, 0.4f) * p.mh04zsd);
                const bool count = p.dfog_on_it == 0 || m.it.it_result == p.dfog_on_it;
                if (s < dstep) {
                    if (count) step_count += (float)(s / dstep);
                    dstep = s;
                } else if (count) {
                    step_count += 1.0f;
                }
                rlast_sw = dstep;
                zz2 += dstep;
                if (zz2 > max_l) {
                    open_air = true;
                    break;
                }
                add_weight(m.it.c, new_vec, dstep);
                zz += dstep * zzplus;
                upd(zz);
                dtmp = m.calc_de();
                if (m.it.it_result >= m.max_its_result) {
                    de_limited = false;
                    break;
                }
                if (dtmp < (double)m.ms_de_stop && m.it.it_result >= p.min_it) break;
                if (dtmp > rlast_de + rlast_sw) dtmp = rlast_de + rlast_sw;
                if (rlast_de > dtmp + 1e-10) {
                    const double d = rlast_sw / (rlast_de - dtmp);
                    rsfmul = d < 1.0 ? (float)std::fmax(d, 0.5) : 1.0f;
                } else {
                    rsfmul = 1.0f;
                }
                if (++guard > 1000000u) {
                    open_air = true;
                    break;
                }
            }
        }
    }
    const double zz2_paint = zz2;
    std::memset(&h, 0, sizeof h);
    const uint16_t shadow = (uint16_t)round_even(std::fmin(std::fmax((double)step_count, 0.0), 1023.0));
    if (!open_air) {
        if (p.de_add_steps != 0) {
            if (de_limited) {
                int itmp = p.de_add_steps;
                double ds = rlast_sw * -0.5;
                double d = m.calc_de();
                while (std::fabs(d - (double)m.ms_de_stop) > 0.001) {
                    zz2 += ds;
                    zz += ds * zzplus;
                    add_weight(m.it.c, new_vec, ds);
                    upd(zz);
                    itmp -= 1;
                    if (itmp <= 0) break;
                    d = m.calc_de();
                    ds = d < (double)m.ms_de_stop ? std::fabs(ds) * -0.55 : std::fabs(ds) * 0.55;
                }
            } else {
                const double before = zz2;
                double vf[3] = {m.vfov[0], m.vfov[1], m.vfov[2]};
                for (int k = 0; k < 3; k++) m.vfov[k] = new_vec[k];
                m.mzz = zz2;
                m.bin_search_it();
                zz2 = m.mzz;
                for (int k = 0; k < 3; k++) m.vfov[k] = vf[k];
                zz += (zz2 - before) * zzplus;
            }
        }
        m.mzz = std::fabs(zz);
        double n[3];
        const float nn = m.calculate_normals(n);
        h.iters = m.nrm_it;
        Px si;
        Marcher::store_normal(si.normal, n);
        float v;
        if (de_limited)
            v = 32767.0f - (nn + p.d_col_plus + p.col_var_de_stop_mul * (float)std::log((double)m.ms_de_stop * sw)) * p.mcts_m;
        else
            v = 32767.0f - nn * p.mcts_m;
        si.si_gradient = clip15(v);
        do_color(m, si);
        double ds = zz2;
        if (depth_sqrt) {
            const double a = zz2 + (double)last_zz;
            ds = std::sqrt(std::fmax(a * a - (double)last_zz + 1e-30, 0.0));
        }
        const int64_t vz = (int64_t)round_even(p.zc_mul * (std::sqrt(std::fmax(ds, 0.0) * p.zcorr + 1.0) - 1.0));
        const int64_t z = std::min<int64_t>(std::max<int64_t>(8388352 - vz, 0), 8388352);
        si.zpos_fine = (uint32_t)z << 8;
        si.shadow = shadow;
        si.amb_shadow = 5000;
        to_words(si, h.gbuf);
    } else {
        h.gbuf[1] = (uint32_t)shadow << 16;
        h.gbuf[2] = 32768u << 16;
    }
    h.flags = (open_air ? 0 : MB3D_HIP_RAY_HIT) | (de_limited ? MB3D_HIP_RAY_DE_LIMITED : 0);
    for (int k = 0; k < 3; k++) h.pos[k] = m.it.c[k];
    h.zz = zz;
    h.zz2 = zz2;
    h.zz2_paint = zz2_paint;
    h.step_count = step_count;
    if (open_air) h.iters = m.it.it_result;
}

// calc_hs (src/reflect.rs:805-952) at a ray's hit: m at the hit (it.c, mzz
// = |zz|, vfov the pixel's view vector), the hit's shadow word in and out
static uint16_t ray_hard_shadows(Marcher &m, const mb3d_hip_ray_opts &o, uint16_t shadow, const int16_t nq[3]) {
    const mb3d_hip_params &p = m.p;
    const double sw = p.step_width;
    const float zz = (float)std::fabs(m.mzz);
    double ic[3] = {m.it.c[0], m.it.c[1], m.it.c[2]};
    const double max_lhs = (double)((float)(p.width + p.height) * 0.6f *
                                    (1.0f + 0.5f * std::fmin(zz, (float)p.zend * 0.4f) * std::fmax((float)p.fov_y, 0.0f) /
                                                (float)p.height) *
                                    o.hs_mul);
    const double nn[3] = {(double)nq[0] / 32767.0, (double)nq[1] / 32767.0, (double)nq[2] / 32767.0};
    double nvec[3];
    rot_rev(nn, p.vgrads, nvec);
    auto upd = [&](double z) { m.ms_de_stop = (float)((double)p.de_stop * (1.0 + std::fabs(z) * (double)p.de_stop_factor)); };
    // the march towards the light (the closure `march`): the soft shadow's
    // value, dt1 the length left
    auto march = [&](const double hsvec[3], double &dt1, double &zz2, double zz2mul, bool soft, float zrs_mul) -> float {
        float rsfd = 1.0f, zrsoft = 1.0f;
        const double dmaxl = dt1;
        double d = m.calc_de();
        int guard = 0;
        for (;;) {
            guard += 1;
            const double last_de = d;
            const double st = (double)std::fmin((float)(d * (double)p.z_step_div * (double)rsfd),
                                                std::fmax(m.ms_de_stop, 0.4f) * p.mh04zsd);
            dt1 -= st;
            add_weight(m.it.c, hsvec, -st);
            zz2 += st * zz2mul;
            upd(zz2);
            d = m.calc_de();
            if (soft) {
                const float r = (float)((dmaxl - dt1) / max_lhs);
                const float r8 = (r * r) * (r * r);
                zrsoft = std::fmin(zrsoft, (float)((d - (double)m.ms_de_stop) * (double)zrs_mul / (dmaxl - dt1 + 0.11)) + r8 * r8);
            }
            if (m.it.it_result >= m.max_its_result || d < (double)m.ms_de_stop) break;
            if (d > last_de + st) d = last_de + st;
            if (last_de > d + 1e-30) {
                const float t = (float)(st / (last_de - d));
                rsfd = t < 1.0f ? std::fmax(t, 0.5f) : 1.0f;
            } else {
                rsfd = 1.0f;
            }
            if (dt1 < 0.0 || guard > 1000000) break;
        }
        return zrsoft;
    };
    auto light_dir = [&](int i, double &len, double hsvec[3]) -> bool {  // false: out of the light's reach
        for (int k = 0; k < 3; k++) hsvec[k] = o.light_vec[i][k];
        if (!o.light_positional[i]) return true;
        double v[3];
        for (int k = 0; k < 3; k++) v[k] = o.light_vec[i][k] - m.it.c[k];
        const double d = dot3(v, v);
        if (d > (double)(o.light_lmax[i] * o.hs_mul)) return false;
        if (d < (len * sw) * (len * sw)) len = std::sqrt(d) / sw;
        const double l = std::sqrt(d);  // norm_to(-sw, v)
        for (int k = 0; k < 3; k++) hsvec[k] = v[k] * (-sw / l);
        return true;
    };
    auto zz2mul_of = [&](const double hsvec[3]) {
        return -dot3(hsvec, m.vfov) / std::sqrt(dot3(hsvec, hsvec) * dot3(m.vfov, m.vfov));
    };
    if (o.soft_radius > 0.0f) {
        const int i = o.n_lights - 1;
        const float sr = std::fmax(o.soft_radius, 0.001f);
        const float zrs_mul = o.light_positional[i] ? ((o.light_visible[i] & 6) == 2 ? 70.0f / sr : 40.0f / sr) : 80.0f / sr;
        shadow |= 0xFC00;
        double zz2 = m.mzz;
        upd((double)zz);
        double dmaxl = max_lhs;
        double hsvec[3];
        if (!light_dir(i, dmaxl, hsvec)) return shadow;
        const double zz2mul = zz2mul_of(hsvec);
        if (dmaxl > 0.0) {
            if (dot3(nvec, hsvec) > 0.0) return shadow & 0x3FF;
            double dt1 = dmaxl;
            const float zr = march(hsvec, dt1, zz2, zz2mul, true, zrs_mul);
            shadow = (uint16_t)((shadow & 0x3FF) | ((uint16_t)round_even((double)(std::fmin(std::fmax(zr, 0.0f), 1.0f) * 63.4f)) << 10));
            for (int k = 0; k < 3; k++) m.it.c[k] = ic[k];
        }
        return shadow;
    }
    uint16_t mask = 0;
    for (int i = 0; i < o.n_lights; i++) mask |= (uint16_t)(0x400 << o.light_idx[i]);
    shadow &= (uint16_t)~mask;
    for (int i = 0; i < o.n_lights; i++) {
        const uint16_t bit = (uint16_t)(0x400 << o.light_idx[i]);
        for (int k = 0; k < 3; k++) m.it.c[k] = ic[k];
        double zz2 = m.mzz;
        upd((double)zz);
        double dt1 = max_lhs;
        double hsvec[3];
        if (!light_dir(i, dt1, hsvec)) {
            shadow |= bit;
            continue;
        }
        const double zz2mul = zz2mul_of(hsvec);
        if (dt1 > 0.0) {
            if (dot3(nvec, hsvec) > 0.0) {
                shadow |= bit;
                continue;
            }
            march(hsvec, dt1, zz2, zz2mul, false, 0.0f);
            if (dt1 > 0.0) shadow |= bit;
        }
    }
    for (int k = 0; k < 3; k++) m.it.c[k] = ic[k];
    return shadow;
}

// deao_1pos (src/reflect.rs:954-1090) at a ray's hit, outside rendering and
// without a random first step: the AO word
static uint16_t ray_deao(Marcher &m, int quality, float max_len, const int16_t nq[3]) {
    const mb3d_hip_params &p = m.p;
    const double sw = p.step_width;
    m.it.calc_sit = false;
    double rot_v[3][3];
    for (int i = 0; i < 3; i++) {  // normalise_matrix_to(sw, vgrads)
        const double d = sw / std::sqrt(dot3(p.vgrads[i], p.vgrads[i]));
        for (int j = 0; j < 3; j++) rot_v[i][j] = p.vgrads[i][j] * d;
    }
    const double s_max_d = (double)max_len * 0.5 * std::sqrt((double)p.height * (double)p.height + (double)p.width * (double)p.width);
    double ms_de_stop = (double)std::fmin((float)((double)p.de_stop * (1.0 + std::fabs(m.mzz) * (double)p.de_stop_factor)), 1e6f);
    double ic[3] = {m.it.c[0], m.it.c[1], m.it.c[2]};
    const double step_ao = ms_de_stop / (double)p.de_stop;
    const double max_dist = s_max_d * std::sqrt(step_ao);
    auto ray_dir = [](double ya, double za, float o[3]) {
        const double sy = std::sin(ya), cy = std::cos(ya), sz = std::sin(za), cz = std::cos(za);
        o[0] = (float)(sy * cz);
        o[1] = (float)(-sy * sz);
        o[2] = (float)(-cy);
    };
    const double tau = 6.283185307179586;
    float dirs[16][3];
    int nd = 0;
    double d_step_mul;
    float d_min_a_dif, corr_w;
    if (quality == 0) {
        const double abr = 60.0 * (3.141592653589793 / 180.0);
        for (int i = 0; i < 3; i++) ray_dir(0.5 * abr, (double)i * tau / 3.0, dirs[nd++]);
        d_step_mul = 1.8;
        d_min_a_dif = -1.0f;
        corr_w = 0.3f;
    } else {
        const double abr = 1.5707963267948966 / ((double)quality + 0.9);
        ray_dir(0.0, 0.0, dirs[nd++]);
        for (int iy = 1; iy <= quality; iy++) {
            const int n = (int)round_even(std::sin(0.25) * tau / abr);
            for (int ix = 0; ix < n; ix++) ray_dir(abr * (double)iy, (double)ix * tau / (double)n, dirs[nd++]);
        }
        d_step_mul = 1.0 + std::sin(abr);
        d_min_a_dif = (float)std::cos(abr * 1.2);
        corr_w = quality == 1 ? 0.2f : 0.1666f;
    }
    float n[3] = {(float)nq[0], (float)nq[1], (float)nq[2]};
    const float l = std::fmax(std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]), 1e-30f);
    for (int k = 0; k < 3; k++) n[k] = n[k] / l;
    float rot_w[3][3];
    normal_matrix(n, rot_w);
    if (p.de_stop_factor != 0.0f)
        ms_de_stop /= d_step_mul * d_step_mul;
    else
        ms_de_stop = (double)p.de_stop / (d_step_mul * d_step_mul);
    const int ray_count = nd;
    const double de_mul = std::sqrt((double)ray_count * 0.5);
    const double abr = 1.2 / std::asin(std::fmin(std::fmax(1.0 / de_mul, -1.0), 1.0));
    const double md_d10 = 0.1 / (max_dist * de_mul);
    float min_ra[16];
    for (int k = 0; k < ray_count; k++) {
        const float *d = dirs[k];
        const float v[3] = {rot_w[0][0] * d[0] + rot_w[0][1] * d[1] + rot_w[0][2] * d[2],
                            rot_w[1][0] * d[0] + rot_w[1][1] * d[1] + rot_w[1][2] * d[2],
                            rot_w[2][0] * d[0] + rot_w[2][1] * d[1] + rot_w[2][2] * d[2]};
        const double vd[3] = {(double)v[0], (double)v[1], (double)v[2]};
        double sv[3];
        rot_rev(vd, rot_v, sv);
        double dt1 = step_ao * d_step_mul;
        double s = 1.0;
        bool end = false;
        int guard = 0;
        for (;;) {
            guard += 1;
            if (dt1 > max_dist) {
                dt1 = max_dist;
                end = true;
            }
            for (int q = 0; q < 3; q++) m.it.c[q] = ic[q];
            add_weight(m.it.c, sv, dt1);
            const double d2 = m.calc_de();
            s = std::fmin(s, (d2 - ms_de_stop + dt1 * md_d10) / dt1);
            if (s < 0.02) break;
            dt1 += std::fmax(d2, dt1 * d_step_mul);
            if (end || guard > 100000) break;
        }
        min_ra[k] = (float)(std::fmax(s, 0.0) * de_mul);
    }
    double amount = 0.0;
    for (int k = 0; k < ray_count; k++) {
        float s_add = 0.0f;
        if (min_ra[k] < 1.0f) {
            const float max_add = 1.0f - min_ra[k];
            for (int k2 = 0; k2 < ray_count; k2++) {
                if (k2 == k) continue;
                const float dt = dirs[k][0] * dirs[k2][0] + dirs[k][1] * dirs[k2][1] + dirs[k][2] * dirs[k2][2];
                if (dt > d_min_a_dif) {
                    // acos of the f32 dot (f32::acos), widened
                    const float ac = std::acos(std::fmin(std::fmax(dt, -1.0f), 1.0f));
                    const float overlap = min_ra[k2] - (float)((double)ac * abr) + 1.0f;
                    if (overlap > 0.0f) s_add += std::fmin(max_add, overlap) * corr_w;
                }
            }
        }
        amount += (double)std::fmin(s_add + min_ra[k], 1.0f);
    }
    for (int q = 0; q < 3; q++) m.it.c[q] = ic[q];
    return (uint16_t)std::fmax(round_even(16383.0 * (1.0 - amount / (double)ray_count)), 0.0);
}

extern "C" int32_t orc_march_rays(const mb3d_hip_params *pp, const mb3d_hip_ray_task *tasks, int64_t n,
                                  int32_t depth_sqrt, const mb3d_hip_ray_opts *opts, mb3d_hip_ray_hit *hits,
                                  int32_t threads) {
    if (int e = check_params(pp)) return e;
    const int64_t chunk = 64;
    const int64_t nchunks = (n + chunk - 1) / chunk;
    for_rows_interleaved((int)nchunks, nthreads(threads), [&](int c) {
        Marcher m(*pp);
        for (int64_t i = (int64_t)c * chunk; i < std::min(n, (int64_t)(c + 1) * chunk); i++) {
            mb3d_hip_ray_hit &h = hits[i];
            march_ray_task(m, tasks[i], depth_sqrt != 0, h);
            if (!opts || !(h.flags & MB3D_HIP_RAY_HIT)) continue;
            // the hit's pass (src/reflect.rs:1647-1670): m at the hit, its DE stop kept
            const float saved = m.ms_de_stop;
            for (int k = 0; k < 3; k++) m.it.c[k] = h.pos[k];
            Px si;
            from_words(h.gbuf, si);
            if (opts->n_lights > 0) {
                m.mzz = std::fabs(h.zz);
                si.shadow = ray_hard_shadows(m, *opts, si.shadow, si.normal);
            }
            if (opts->ao) {
                m.mzz = std::fabs(h.zz);
                si.amb_shadow = ray_deao(m, tasks[i].ao_q, opts->ao_max_len, si.normal);
            }
            m.ms_de_stop = saved;
            to_words(si, h.gbuf);
        }
    });
    return 0;
}

// ---------------------------------------------------------------------------
// the Monte Carlo renderer's marches: McMarch::run (src/mc.rs:488-977) for
// this backend's scenes -- outside rendering without the inside-and-outside
// turn-around, no volumetric light, no ground plane, no dIFS parts, colour
// option 0, normals on the DE.

// GetRand (rand_of, src/mc.rs:470-476)
static double mc_rand(int32_t &seed) {
    seed = (int32_t)((uint32_t)seed * 0x343FDu + 0x269EC3u);
    return (double)(seed & 0x7FFFFFFF) * (1.0 / 2147483647.0);
}

// not_zero (src/mc.rs:53-63)
static double mc_not_zero(double d) {
    if (std::fabs(d) < 1e-40) return d < 0.0 ? -1e-40 : 1e-40;
    return d;
}

// update_de_stop of McMarch (src/mc.rs:504-508): the DE stop at view depth z
static void mc_de_stop(Marcher &m, double z) {
    m.ms_de_stop = (float)((double)m.p.de_stop * (1.0 + std::fabs(z) * (double)m.p.de_stop_factor));
}

// McMarch::normals with normals on the DE (src/mc.rs:531-590): the one-sided
// gradient on the world axes z, x, y, its side from bit 7 of the seed, then
// random offsets until it faces the ray (four tries); the smoothed
// iteration count of the centre is returned
static float mc_normals(Marcher &m, const double raydir[3], int32_t &seed, double out[3]) {
    const mb3d_hip_params &p = m.p;
    float sca = (seed & 128) == 0 ? 0.15f : -0.15f;
    m.it.calc_sit = true;
    const double no = m.calc_de();
    const float nn = m.it.smooth_it;
    m.it.calc_sit = false;
    int i = 4;
    double n[3] = {0, 0, 0};
    bool found = false;
    for (;;) {
        if (i < 4) sca = (float)((mc_rand(seed) - 0.5) * 0.5);
        const double off = (double)std::fmin(p.de_stop, 1.0f) * (1.0 + m.mzz * (double)p.de_stop_factor) * (double)sca *
                           p.step_width;
        for (int k : {2, 0, 1}) {
            const double t = m.it.c[k];
            m.it.c[k] = t + off;
            n[k] = (m.calc_de() - no) * (double)sca;
            m.it.c[k] = t;
        }
        const double l = n[0] * n[0] + n[1] * n[1] + n[2] * n[2];
        if (l > 1e-100 && dot3(n, raydir) < 0.0) {
            const double d = 1.0 / std::sqrt(l);
            for (int k = 0; k < 3; k++) out[k] = n[k] * d;
            found = true;
            break;
        }
        i -= 1;
        if (i == 0) break;
    }
    if (!found) {  // norm_to(-1, raydir)
        const double d = -1.0 / std::sqrt(dot3(raydir, raydir) + 1e-100);
        for (int k = 0; k < 3; k++) out[k] = raydir[k] * d;
    }
    return nn;
}

// McMarch::bin_search_it (src/mc.rs:592-643): ten turns on the ray from
// depth zz, through the Marcher's bin_search_it with that depth and ray
static void mc_bin_search_it(Marcher &m, double &zz, const double dir[3]) {
    double vf[3] = {m.vfov[0], m.vfov[1], m.vfov[2]};
    for (int k = 0; k < 3; k++) m.vfov[k] = dir[k];
    const double mzz = m.mzz;
    m.mzz = zz;
    m.bin_search_it(10);
    zz = m.mzz;
    m.mzz = mzz;
    for (int k = 0; k < 3; k++) m.vfov[k] = vf[k];
}

// surface_colour (src/mc.rs:645-662) with colour option 0: gradient, orbit trap
static void mc_colour(Marcher &m, float nn, bool de_limited, mb3d_hip_mc_res &o) {
    const mb3d_hip_params &p = m.p;
    float v;
    if (de_limited)
        v = 32767.0f - (nn + p.d_col_plus + p.col_var_de_stop_mul * (float)std::log((double)m.ms_de_stop * p.step_width)) *
                           p.mcts_m;
    else
        v = 32767.0f - nn * p.mcts_m;
    o.si_grad = clip15(v);
    o.otrap = clip15((float)(m.it.otrap * 4096.0));
}

// McMarch::ray (src/mc.rs:664-866): CalcRay's march
static void mc_ray(Marcher &m, const mb3d_hip_mc_task &r, bool secant, bool inside, mb3d_hip_mc_res &o) {
    const mb3d_hip_params &p = m.p;
    int32_t seed = r.seed;
    const double *dir = r.dir;
    double zz = r.zz, zz2 = r.zz2;
    const double max_l = r.len, zzplus = r.zzplus;
    mc_de_stop(m, zz);
    double dtmp = m.calc_de();
    if ((r.flags & MB3D_HIP_MC_PROBE) && (m.it.it_result >= m.max_its_result || dtmp < (double)m.ms_de_stop)) {
        // the start is inside the set (src/mc.rs:691-727; no inside-and-outside turn-around)
        o.otrap = clip15((float)(m.it.otrap * 4096.0));
        m.mzz = zz;
        if (inside) {
            m.it.calc_sit = true;
            m.calc_de();
            o.si_grad = (uint16_t)(clip15(m.it.smooth_it * p.mcts_m) | 32768u);
        } else {
            const double t = std::fmin(std::fmax(m.it.rout / p.d_rstop, 0.0), 1.0);
            o.si_grad = (uint16_t)(32768u + (uint32_t)round_even(32767.0 * t));
        }
        o.kind = 2;
        for (int k = 0; k < 3; k++) o.pos[k] = m.it.c[k];
        o.flags = inside ? MB3D_HIP_MC_INSIDE : 0;
        o.seed = seed;
        return;
    }
    float dyn_fog = 0.0f, rsfmul = 1.0f;
    bool open_air = false, de_limited = true, first_step = true;
    double rlast_sw = dtmp * (double)p.z_step_div, rlast_de = 0.0;
    uint32_t guard = 0;
    for (;;) {
        rlast_de = dtmp;
        double dstep = std::fmax((dtmp - (double)p.ms_de_sub * (double)m.ms_de_stop) * (double)p.z_step_div * (double)rsfmul, 0.11);
        const double d1 = (double)(std::fmax(m.ms_de_stop, 0.4f) * p.mh04zsd);
        const bool count = p.dfog_on_it == 0 || m.it.it_result == p.dfog_on_it;
        if (d1 < dstep) {
            if (count) dyn_fog += (float)(d1 / dstep);
            dstep = d1;
        } else if (count) {
            dyn_fog += 1.0f;
        }
        if (first_step) {
            first_step = false;
            dstep *= mc_rand(seed);
        }
        rlast_sw = dstep;
        zz2 += dstep;
        if (zz2 > max_l) {
            open_air = true;
            break;
        }
        add_weight(m.it.c, dir, dstep);
        zz += dstep * zzplus;
        mc_de_stop(m, zz);
        dtmp = m.calc_de();
        if (m.it.it_result >= m.max_its_result) {
            de_limited = false;
            break;
        }
        if (!(dtmp >= (double)m.ms_de_stop) && m.it.it_result >= p.min_it) break;
        if (dtmp > rlast_de + rlast_sw) dtmp = rlast_de + rlast_sw;
        if (rlast_de > dtmp + 1e-30) {
            const double d = rlast_sw / (rlast_de - dtmp);
            rsfmul = d < 1.0 ? (float)std::fmax(d, 0.5) : 1.0f;
        } else {
            rsfmul = 1.0f;
        }
        if (++guard > 1000000u) {
            open_air = true;
            break;
        }
    }
    if (inside) dyn_fog *= 200.0f * p.de_stop / (float)p.width;
    o.dyn_fog = dyn_fog;
    o.zz2 = zz2;
    o.flags = inside ? MB3D_HIP_MC_INSIDE : 0;
    if (open_air) {
        for (int k = 0; k < 3; k++) o.pos[k] = m.it.c[k];
        o.zz = zz;
        o.seed = seed;
        return;
    }
    if (de_limited) {
        if (secant) {  // src/mc.rs:806-830
            int itmp = 4;
            while (itmp > 0 && std::fabs(dtmp - (double)m.ms_de_stop) > 0.001) {
                rlast_de = mc_not_zero(rlast_de - dtmp);
                const double ms = (double)m.ms_de_stop;
                const double f = rlast_sw * (dtmp - ms) / rlast_de;
                if (dtmp < ms)
                    rlast_sw = (f >= 0.0 || f < std::fabs(rlast_sw) * -0.94) ? std::fabs(rlast_sw) * -0.5 : f;
                else
                    rlast_sw = (f <= 0.0 || f > std::fabs(rlast_sw) * 0.94) ? std::fabs(rlast_sw) * 0.5 : f;
                rlast_de = dtmp;
                zz += rlast_sw * zzplus;
                add_weight(m.it.c, dir, rlast_sw);
                mc_de_stop(m, zz);
                itmp -= 1;
                if (itmp <= 0) break;
                dtmp = m.calc_de();
            }
        } else {  // src/mc.rs:831-846
            int itmp = 10;
            double ds = rlast_sw * -0.5;
            while (std::fabs(dtmp - (double)m.ms_de_stop) > 0.001) {
                zz += ds * zzplus;
                add_weight(m.it.c, dir, ds);
                mc_de_stop(m, zz);
                itmp -= 1;
                if (itmp <= 0) break;
                dtmp = m.calc_de();
                ds = dtmp < (double)m.ms_de_stop ? std::fabs(ds) * -0.55 : std::fabs(ds) * 0.55;
            }
        }
    } else {
        double ds = zz;
        mc_bin_search_it(m, ds, dir);
        zz += (ds - zz) * zzplus;
    }
    m.mzz = std::fabs(zz);
    double n[3];
    const float nn = mc_normals(m, dir, seed, n);
    for (int k = 0; k < 3; k++) o.pos[k] = m.it.c[k];
    mc_colour(m, nn, de_limited, o);
    o.kind = 1;
    o.zz = zz;
    if (de_limited) o.flags |= MB3D_HIP_MC_DE_LIMITED;
    for (int k = 0; k < 3; k++) o.normals[k] = n[k];
    o.seed = seed;
}

// McMarch::shadow (src/mc.rs:868-956): CalcHSMC's ray towards a light
static void mc_shadow(Marcher &m, const mb3d_hip_mc_task &r, mb3d_hip_mc_res &o) {
    const mb3d_hip_params &p = m.p;
    int32_t seed = r.seed;
    const double *v = r.dir;
    double dt1 = r.len, zz2 = r.zz2;
    const double zz2mul = r.zzplus;
    mc_de_stop(m, r.zz);
    float rsfd = 2.0f;
    double dtmp = m.calc_de(), rlast_de = 0.0, rlast_sw = 0.0;
    for (;;) {
        rlast_de = dtmp;
        dtmp = std::fmin(dtmp * (double)p.z_step_div * (double)rsfd, (double)(std::fmax(m.ms_de_stop, 0.4f) * p.mh04zsd));
        rlast_sw = dtmp;
        dt1 -= dtmp;
        add_weight(m.it.c, v, dtmp);
        zz2 += dtmp * zz2mul;
        mc_de_stop(m, zz2);
        dtmp = m.calc_de();
        if (m.it.it_result >= m.max_its_result || !(dtmp >= (double)m.ms_de_stop) || !(dtmp < 1e30)) break;
        if (dtmp > rlast_de + rlast_sw) dtmp = rlast_de + rlast_sw;
        if (rlast_de > dtmp + 1e-30) {
            const double s = rlast_sw / (rlast_de - dtmp);
            rsfd = s < 1.0 ? (float)std::fmax(s, 0.5) : 1.0f;
        } else {
            rsfd = 1.0f;
        }
        if (!(dt1 >= 0.0)) break;
    }
    o.len = dt1;
    o.seed = seed;
    if (dt1 <= 0.01 || !(r.flags & MB3D_HIP_MC_TRANS)) return;
    // a transparent object between the surface and the light
    if (m.it.it_result < m.max_its_result) {
        int itmp = 3;
        while (itmp > 0 && std::fabs(dtmp - (double)m.ms_de_stop) > 0.01) {
            rlast_de = mc_not_zero(rlast_de - dtmp);
            const double ms = (double)m.ms_de_stop;
            const double f = rlast_sw * (dtmp - ms) / rlast_de;
            if (dtmp < ms)
                rlast_sw = (f >= 0.0 || f < std::fabs(rlast_sw) * -0.94) ? std::fabs(rlast_sw) * -0.5 : f;
            else
                rlast_sw = (f <= 0.0 || f > std::fabs(rlast_sw) * 0.94) ? std::fabs(rlast_sw) * 0.5 : f;
            rlast_de = dtmp;
            zz2 += rlast_sw * zz2mul;
            add_weight(m.it.c, v, rlast_sw);
            mc_de_stop(m, zz2);
            dtmp = m.calc_de();
            itmp -= 1;
        }
    } else {
        double d = zz2;
        mc_bin_search_it(m, d, v);
        zz2 += (d - zz2) * zz2mul;
    }
    m.mzz = std::fabs(zz2);
    double n[3];
    const float nn = mc_normals(m, v, seed, n);
    for (int k = 0; k < 3; k++) o.pos[k] = m.it.c[k];
    // the iteration count after the normals, as in MB3D's CalcHSMC
    mc_colour(m, nn, m.it.it_result < m.max_its_result, o);
    o.kind = 1;
    o.zz2 = zz2;
    for (int k = 0; k < 3; k++) o.normals[k] = n[k];
    o.ms_de_stop = m.ms_de_stop;
    o.seed = seed;
}

// McMarch::through (src/mc.rs:958-977): from the far end back towards the
// start, in steps of at least one, until the DE stop or the start
static void mc_through(Marcher &m, const mb3d_hip_mc_task &r, mb3d_hip_mc_res &o) {
    int32_t seed = r.seed;
    const double *v = r.dir;
    double dt1 = r.len;
    m.ms_de_stop = r.ms_de_stop;
    add_weight(m.it.c, v, dt1);
    double d = mc_rand(seed) * 8.0 * std::fmax((double)m.ms_de_stop, 1.0);
    for (;;) {
        d = -1.0 - std::fabs(d);
        dt1 += d;
        add_weight(m.it.c, v, d);
        d = m.calc_de();
        if (d > (double)m.ms_de_stop || !(dt1 > 0.0)) break;
    }
    o.len = dt1;
    o.seed = seed;
}

// McMarch::run (src/mc.rs:488-502) for one task, on a fresh marcher
static void mc_run(const mb3d_hip_params &p, const mb3d_hip_mc_task &r, bool secant, mb3d_hip_mc_res &o) {
    Marcher m(p);
    std::memset(&o, 0, sizeof o);
    for (int k = 0; k < 3; k++) m.it.c[k] = r.pos[k];
    const bool inside = (r.flags & MB3D_HIP_MC_INSIDE) != 0;
    m.calc_inside = inside;
    m.it.calc_sit = false;
    if (r.kind == MB3D_HIP_MC_RAY)
        mc_ray(m, r, secant, inside, o);
    else if (r.kind == MB3D_HIP_MC_SHADOW)
        mc_shadow(m, r, o);
    else
        mc_through(m, r, o);
}

extern "C" int32_t orc_mc_batch(const mb3d_hip_params *pp, const mb3d_hip_mc_task *tasks, int64_t n, int32_t secant,
                                mb3d_hip_mc_res *res, int32_t threads) {
    if (int e = check_params(pp)) return e;
    if (n < 0 || (n > 0 && (!tasks || !res))) return fail("bad argument");
    for (int64_t i = 0; i < n; i++)
        if (tasks[i].kind < MB3D_HIP_MC_RAY || tasks[i].kind > MB3D_HIP_MC_THROUGH) return fail("task kind not 0..2");
    const int64_t chunk = 64;
    const int64_t nchunks = (n + chunk - 1) / chunk;
    for_rows_interleaved((int)nchunks, nthreads(threads), [&](int c) {
        for (int64_t i = (int64_t)c * chunk; i < std::min(n, (int64_t)(c + 1) * chunk); i++)
            mc_run(*pp, tasks[i], secant != 0, res[i]);
    });
    return 0;
}

namespace {
// the cone grid of a rendering (orc_render_cone): blocks of bw x bh pixels
struct ConeGrid {
    int bw = 0, bh = 0, cols = 0, mode = 0;
    std::vector<ConeCell> cells;
    const ConeCell *at(int x, int y) const { return bw ? &cells[(size_t)(y / bh) * cols + x / bw] : nullptr; }
};

ConeGrid cone_grid(const mb3d_hip_params &p, int bw, int bh, int mode, int threads) {
    ConeGrid g;
    g.bw = bw;
    g.bh = bh;
    g.mode = mode;
    g.cols = (p.width + bw - 1) / bw;
    const int rows = (p.height + bh - 1) / bh;
    g.cells.resize((size_t)g.cols * rows);
    for_rows_interleaved(rows, nthreads(threads), [&](int by) {
        Marcher m(p);
        for (int bx = 0; bx < g.cols; bx++) g.cells[(size_t)by * g.cols + bx] = cone_march(m, bx, by, bw, bh);
    });
    return g;
}

int render_impl(const mb3d_hip_params *pp, const mb3d_hip_post_job *post, int32_t row0, int32_t rows,
                const orc_out *out, int32_t threads, orc_counters *cnt, const ConeGrid *cone);
}  // namespace

extern "C" int32_t orc_render(const mb3d_hip_params *pp, const mb3d_hip_post_job *post, int32_t row0, int32_t rows,
                              const orc_out *out, int32_t threads, orc_counters *cnt) {
    return render_impl(pp, post, row0, rows, out, threads, cnt, nullptr);
}

// One pixel's march with what do_color saw at its surface: out = the
// orbit's end v (3), the surface point c (3), rout, rold; words = the
// otrap and si_gradient words.  Returns 1 at a surface, 0 otherwise.
extern "C" int32_t orc_color_probe(const mb3d_hip_params *pp, int32_t x, int32_t y, double *out, uint16_t *words) {
    if (int e = check_params(pp)) return e;
    if (x < 0 || y < 0 || x >= pp->width || y >= pp->height) return fail("pixel out of the picture");
    Marcher m(*pp);
    m.seed = orc_pixel_seed(y, x);
    ColorProbe pr;
    std::memset(&pr, 0, sizeof pr);
    g_color_probe = &pr;
    Px s = march_pixel(m, x, y);
    g_color_probe = nullptr;
    if (!pr.set || !(std::isfinite(s.depth) && s.depth > 0)) return 0;
    for (int k = 0; k < 3; k++) {
        out[k] = pr.v[k];
        out[3 + k] = pr.c[k];
    }
    out[6] = pr.rout;
    out[7] = pr.rold;
    words[0] = pr.otrap;
    words[1] = pr.gradient;
    return 1;
}

extern "C" int32_t orc_cone_grid(const mb3d_hip_params *pp, int32_t bw, int32_t bh, double *t, float *count,
                                 int32_t threads) {
    if (int e = check_params(pp)) return e;
    if (bw < 1 || bh < 1) return fail("bad block size");
    ConeGrid g = cone_grid(*pp, bw, bh, 1, threads);
    for (size_t i = 0; i < g.cells.size(); i++) {
        t[i] = g.cells[i].t;
        count[i] = g.cells[i].count;
    }
    return 0;
}

extern "C" int32_t orc_render_cone(const mb3d_hip_params *pp, int32_t bw, int32_t bh, int32_t mode, int32_t row0,
                                   int32_t rows, const orc_out *out, int32_t threads, orc_counters *cnt) {
    if (int e = check_params(pp)) return e;
    if (bw < 1 || bh < 1 || mode < 1 || mode > 2) return fail("bad cone block or mode");
    ConeGrid g = cone_grid(*pp, bw, bh, mode, threads);
    return render_impl(pp, nullptr, row0, rows, out, threads, cnt, &g);
}

namespace {
int render_impl(const mb3d_hip_params *pp, const mb3d_hip_post_job *post, int32_t row0, int32_t rows,
                const orc_out *out, int32_t threads, orc_counters *cnt, const ConeGrid *cone) {
    if (int e = check_params(pp)) return e;
    if (!out) return fail("out is NULL");
    const mb3d_hip_params &p = *pp;
    if (row0 < 0 || rows < 0 || row0 + rows > p.height) return fail("rows out of the picture");
    int w = p.width;
    std::vector<Px> px((size_t)rows * w);
    std::atomic<uint64_t> c_steps{0}, c_de{0}, c_p8{0}, c_box{0}, c_oth{0}, c_hits{0};
    for_rows_interleaved(rows, nthreads(threads), [&](int r) {
        Marcher m(p);
        int y = row0 + r;
        uint64_t hits = 0, steps = 0;
        for (int x = 0; x < w; x++) {
            m.seed = orc_pixel_seed(y, x);
            Px s = cone ? march_pixel(m, x, y, cone->at(x, y), cone->mode) : march_pixel(m, x, y);
            hits += std::isfinite(s.depth) && s.depth > 0 ? 1 : 0;
            steps += (uint64_t)s.steps;
            px[(size_t)r * w + x] = s;
        }
        c_steps += steps;
        c_hits += hits;
        c_de += m.de_evals;
        c_p8 += m.it.its_p8;
        c_box += m.it.its_box;
        c_oth += m.it.its_other;
    });
    if (post) {
        if (int e = check_job(*post)) return e;
        run_post(p, *post, row0, rows, px, threads);
    }
    for (size_t i = 0; i < px.size(); i++) {
        const Px &s = px[i];
        if (out->depth) out->depth[i] = s.depth;
        if (out->iters) out->iters[i] = s.iters;
        if (out->normals)
            for (int k = 0; k < 3; k++) out->normals[3 * i + k] = s.n[k];
        if (out->gbuf) to_words(s, out->gbuf + 5 * i);
        if (out->steps) out->steps[i] = s.steps;
    }
    if (cnt) {
        cnt->pixels = px.size();
        cnt->hits = c_hits;
        cnt->steps = c_steps;
        cnt->de_evals = c_de;
        cnt->its_p8 = c_p8;
        cnt->its_box = c_box;
        cnt->its_other = c_oth;
    }
    return 0;
}

}  // namespace

extern "C" int32_t orc_post(const mb3d_hip_params *pp, const mb3d_hip_post_job *post, int32_t row0, int32_t rows,
                            uint32_t *gbuf, int32_t threads) {
    if (int e = check_params(pp)) return e;
    if (!post || !gbuf) return fail("NULL argument");
    if (int e = check_job(*post)) return e;
    const mb3d_hip_params &p = *pp;
    int w = p.width;
    std::vector<Px> px((size_t)rows * w);
    for (size_t i = 0; i < px.size(); i++) from_words(gbuf + 5 * i, px[i]);
    run_post(p, *post, row0, rows, px, threads);
    for (size_t i = 0; i < px.size(); i++) to_words(px[i], gbuf + 5 * i);
    return 0;
}
