// params.cpp -- host side of the HIP backend: scene defaults, the derived
// march parameters (the reference's CalcParams::new for the built-in
// formulas), and the backend's capability check.
//
//   mb3d_hip_scene_default      Scene::default          src/scene.rs:626-700
//   mb3d_hip_formula_default    Formula::default_for    src/formulas.rs:266-306
//   mb3d_hip_params_from_scene  CalcParams::new         src/calc.rs:375-898
//   mb3d_hip_supported          gpu::unsupported        src/gpu.rs:445-477
//   mb3d_hip_pixel_seed         pixel_seed              src/calc.rs:1125-1140
#include <algorithm>
#include <cctype>
#include <cmath>
#include <initializer_list>
#include <cstdio>
#include <cstring>
#include <string>

#include "../../include/mb3d_hip.h"
#include "host_common.h"

namespace {

constexpr double kPi = 3.141592653589793;
constexpr double kRadPerDeg = kPi / 180.0;  // f64::to_radians

// The built-in formulas' metadata (src/formulas.rs:492-560): DE option,
// DE scale, default bailout, smooth-iteration power, first option.
struct FormulaInfo {
    int de_option;
    double de_scale;
    double default_rstop;
    double si_pow;
    double first_option;
    bool first_option_double;
};

FormulaInfo info_of(const mb3d_hip_formula &f) {
    const double *o = f.options;
    // a real power of 0 counts as 1e-40 for the smooth iterations
    const double real_sip = o[0] == 0.0 ? 1e-40 : o[0];
    switch (f.kind) {
        case MB3D_HIP_FORMULA_AMAZINGBOX: return {11, 0.2, 1024.0, 2.0, f.scale, true};
        case MB3D_HIP_FORMULA_REALPOWER: return {0, 1.0, 16.0, real_sip, o[0], true};
        case MB3D_HIP_FORMULA_QUATERNION: return {4, 1.0, 16.0, 2.0, o[0], true};
        case MB3D_HIP_FORMULA_TRICORN: return {0, 1.0, 16.0, 2.0, o[0], true};
        case MB3D_HIP_FORMULA_BULBOX: return {0, 0.2, 1024.0, 2.0, o[0], true};
        case MB3D_HIP_FORMULA_FOLDINGINTPOW: return {0, 0.5, 1024.0, 2.0, o[1], false};
        case MB3D_HIP_FORMULA_AEXIONC: return {0, 1.0, 16.0, real_sip, o[0], true};
        default: return {0, 1.0, 16.0, (double)std::min(std::max(f.power, 2), 8), f.z_mul, false};
    }
}

// combine_de_options (src/calc.rs:318-372): the DE option of the hybrid so
// far (`acc`, -1 at first) and that of the next formula
int combine_de(int acc, int de) {
    auto in = [de](std::initializer_list<int> l) { return std::find(l.begin(), l.end(), de) != l.end(); };
    switch (acc) {
        case -1:
        case 21:
        case 22: return de;
        case 2: return in({2, 5, 6, 11}) ? 2 : 0;
        case 4: return in({5, 6}) ? 4 : 0;
        case 5: return de == 4 ? 4 : in({2, 11}) ? 2 : (de != 6 ? 0 : 5);
        case 6: return in({2, 4, 5, 11}) ? de : 0;
        case 11: return in({2, 5}) ? 2 : (de != 6 ? 0 : 11);
        default: return acc;
    }
}

// An integer option as Formula::set_option stores it: rounded half away
// from zero, clamped (NaN casts to 0 and is clamped likewise),
// src/formulas.rs:383-404
int int_option(double v, int lo, int hi) {
    const double r = std::round(v);
    if (!(r >= lo)) return lo;
    return r > hi ? hi : (int)r;
}

// The options of the formula kinds that keep them in `options`, as the
// formula holds them: Folding Int Pow's power an integer, Aexion C's
// switches 0/1 and its mode 0..31.  Returns the integer power (or 0).
int slot_options(const mb3d_hip_formula &f, double out[MB3D_HIP_FORMULA_OPTIONS]) {
    std::copy(f.options, f.options + MB3D_HIP_FORMULA_OPTIONS, out);
    if (f.kind == MB3D_HIP_FORMULA_FOLDINGINTPOW) {
        const int pw = int_option(f.options[0], 2, 8);
        out[0] = pw;
        return pw;
    }
    if (f.kind == MB3D_HIP_FORMULA_AEXIONC) {
        for (int k : {2, 3, 6}) out[k] = std::round(f.options[k]) != 0.0 ? 1.0 : 0.0;
        out[7] = int_option(f.options[7], 0, 31);
    }
    return 0;
}

// build_smatrix4 (src/math.rs:178-204): the XW, YW and ZW plane rotations
// applied in that order (m3 * m2 * m1), each entry rounded to f32
void smatrix4(const double a[3], double out[4][4]) {
    double m[3][4][4];
    for (int r = 0; r < 3; r++) {
        for (int i = 0; i < 4; i++)
            for (int k = 0; k < 4; k++) m[r][i][k] = i == k ? 1.0 : 0.0;
        const double sn = std::sin(a[r]), cs = std::cos(a[r]);
        m[r][r][r] = cs;
        m[r][3][3] = cs;
        m[r][r][3] = -sn;
        m[r][3][r] = sn;
    }
    double acc[4][4], t[4][4];
    std::memcpy(acc, m[0], sizeof acc);
    for (int r = 1; r < 3; r++) {
        for (int i = 0; i < 4; i++)
            for (int k = 0; k < 4; k++) {
                double sum = 0.0;
                for (int q = 0; q < 4; q++) sum += m[r][i][q] * acc[q][k];
                t[i][k] = sum;
            }
        std::memcpy(acc, t, sizeof acc);
    }
    for (int i = 0; i < 4; i++)
        for (int k = 0; k < 4; k++) out[i][k] = (double)(float)acc[i][k];
}

// Formula::default_for's names and aliases (src/formulas.rs:266-306),
// compared lower-case with '_' and '-' read as spaces
int kind_of_name(const char *name) {
    std::string n;
    for (const char *c = name; *c; c++) {
        const char ch = (char)std::tolower((unsigned char)*c);
        n += (ch == '_' || ch == '-') ? ' ' : ch;
    }
    const size_t b = n.find_first_not_of(" \t\r\n"), e = n.find_last_not_of(" \t\r\n");
    n = b == std::string::npos ? std::string() : n.substr(b, e - b + 1);
    static const struct {
        const char *name;
        int kind;
    } table[] = {
        {"integer power", MB3D_HIP_FORMULA_INTPOW}, {"intpow", MB3D_HIP_FORMULA_INTPOW},
        {"mandelbulb", MB3D_HIP_FORMULA_INTPOW}, {"bulb", MB3D_HIP_FORMULA_INTPOW},
        {"real power", MB3D_HIP_FORMULA_REALPOWER}, {"realpower", MB3D_HIP_FORMULA_REALPOWER},
        {"floatpow", MB3D_HIP_FORMULA_REALPOWER}, {"quaternion", MB3D_HIP_FORMULA_QUATERNION},
        {"quat", MB3D_HIP_FORMULA_QUATERNION}, {"tricorn", MB3D_HIP_FORMULA_TRICORN},
        {"amazing box", MB3D_HIP_FORMULA_AMAZINGBOX}, {"amazingbox", MB3D_HIP_FORMULA_AMAZINGBOX},
        {"mandelbox", MB3D_HIP_FORMULA_AMAZINGBOX}, {"box", MB3D_HIP_FORMULA_AMAZINGBOX},
        {"bulbox", MB3D_HIP_FORMULA_BULBOX}, {"folding int pow", MB3D_HIP_FORMULA_FOLDINGINTPOW},
        {"foldingintpow", MB3D_HIP_FORMULA_FOLDINGINTPOW}, {"foldintpow", MB3D_HIP_FORMULA_FOLDINGINTPOW},
        {"aexion c", MB3D_HIP_FORMULA_AEXIONC}, {"aexionc", MB3D_HIP_FORMULA_AEXIONC},
        {"aexion", MB3D_HIP_FORMULA_AEXIONC},
    };
    for (const auto &t : table)
        if (n == t.name) return t.kind;
    return MB3D_HIP_FORMULA_NONE;
}

}  // namespace

extern "C" void mb3d_hip_scene_default(mb3d_hip_scene *s) {
    if (!s) return;
    std::memset(s, 0, sizeof *s);
    s->abi_version = MB3D_HIP_ABI_VERSION;
    s->width = 800;
    s->height = 600;
    s->iterations = 60;
    s->min_iterations = 1;
    s->bin_search_steps = 8;
    s->first_step_random = 1;
    s->vary_de_stop_on_fov = 1;
    s->normals_on_de = 1;
    s->n_formulas = 1;
    s->zoom = 1.0;
    s->z_start = -2.0;
    s->z_end = 30.0;
    for (int i = 0; i < 3; i++) s->vgrads[i][i] = 1.0;
    s->fov_y = 30.0;
    s->de_stop = 1.0;
    s->z_step_div = 0.5;
    s->raystep_limiter = 1.0;
    s->color_mul = 1.0;
    // DeCombParams::new (src/scene.rs:2196-2200) for when a combination is switched on
    s->decomb_end1 = 0;
    s->decomb_start2 = 1;
    s->decomb_end2 = 5;
    s->decomb_repeat2 = 1;
    s->decomb_iterations2 = s->iterations;
    s->decomb_smooth = 0.5f;
    s->part2_scale = 1.0;
    s->formulas[0].kind = MB3D_HIP_FORMULA_INTPOW;
    s->formulas[0].iterations = 1;
    s->formulas[0].power = 8;
    s->formulas[0].z_mul = -1.0;
}

extern "C" int32_t mb3d_hip_formula_default(const char *name, mb3d_hip_formula *f) {
    using mb3d::set_error;
    if (!name || !f) return set_error(MB3D_HIP_ERR_INVALID, "NULL argument");
    const int kind = kind_of_name(name);
    if (kind == MB3D_HIP_FORMULA_NONE) {
        const std::string why = std::string("not a built-in formula: ") + name;
        return set_error(MB3D_HIP_UNSUPPORTED, why.c_str());
    }
    std::memset(f, 0, sizeof *f);
    f->kind = kind;
    f->iterations = 1;
    double *o = f->options;
    switch (kind) {
        case MB3D_HIP_FORMULA_INTPOW:
            f->power = 8;
            f->z_mul = -1.0;
            break;
        case MB3D_HIP_FORMULA_AMAZINGBOX:
            f->scale = 2.0;
            f->min_r = 0.5;
            f->fold = 1.0;
            break;
        case MB3D_HIP_FORMULA_REALPOWER:
            o[0] = 8.0;
            o[1] = -1.0;
            break;
        case MB3D_HIP_FORMULA_QUATERNION:
            o[0] = -1.0;
            o[1] = 0.0;
            break;
        case MB3D_HIP_FORMULA_TRICORN:
            o[0] = -2.0;
            o[1] = 1.0;
            break;
        case MB3D_HIP_FORMULA_BULBOX: {
            const double d[6] = {2.0, 0.5, 1.0, 1.0, 2.0, 2.0};
            std::copy(d, d + 6, o);
            break;
        }
        case MB3D_HIP_FORMULA_FOLDINGINTPOW:
            o[0] = 2.0;
            o[1] = -1.0;
            o[2] = 2.0;
            break;
        case MB3D_HIP_FORMULA_AEXIONC: {
            const double d[8] = {8.0, 1.0, 1.0, 1.0, 8.0, 1.0, 0.0, 0.0};
            std::copy(d, d + 8, o);
            break;
        }
    }
    return MB3D_HIP_OK;
}

extern "C" int32_t mb3d_hip_params_from_scene(const mb3d_hip_scene *sc, mb3d_hip_params *out) {
    using mb3d::set_error;
    if (!sc || !out) return set_error(MB3D_HIP_ERR_INVALID, "NULL argument");
    if (sc->abi_version != MB3D_HIP_ABI_VERSION) return set_error(MB3D_HIP_ERR_INVALID, "scene ABI version mismatch");
    const int nf = sc->n_formulas;
    if (nf < 1 || nf > MB3D_HIP_MAX_SLOTS) return set_error(MB3D_HIP_ERR_INVALID, "a hybrid has 1..6 formulas");
    if (sc->width <= 0 || sc->height <= 0) return set_error(MB3D_HIP_ERR_INVALID, "bad picture size");
    bool any_its = false;
    for (int i = 0; i < nf; i++) {
        const int k = sc->formulas[i].kind;
        if (k < MB3D_HIP_FORMULA_INTPOW || k > MB3D_HIP_FORMULA_AEXIONC)
            return set_error(MB3D_HIP_ERR_INVALID, "unknown formula kind");
        any_its |= sc->formulas[i].iterations != 0;
    }
    // the interpolation hybrid of the first two formulas, its weights
    // normalised in f32 (src/calc.rs:378-397)
    const bool ipol = sc->interpolation != 0;
    float ipw[2] = {1.0f, 0.0f};
    if (ipol) {
        if (nf < 2) return set_error(MB3D_HIP_ERR_INVALID, "an interpolation hybrid needs two formulas");
        const float sum = sc->interpolation_w[0] + sc->interpolation_w[1];
        if (sum < 1e-10f) {
            ipw[0] = 1.0f;
            ipw[1] = 0.0f;
        } else {
            const float inv = 1.0f / sum;
            ipw[0] = sc->interpolation_w[0] * inv;
            ipw[1] = sc->interpolation_w[1] * inv;
        }
    } else if (!any_its) {
        return set_error(MB3D_HIP_ERR_INVALID, "no formula with iterations > 0");
    }

    mb3d_hip_params p;
    std::memset(&p, 0, sizeof p);
    p.abi_version = MB3D_HIP_ABI_VERSION;
    p.width = sc->width;
    p.height = sc->height;
    const double W = sc->width, Hh = sc->height;

    // the hybrid ends at the last slot with iterations (CheckHybridOptions)
    // (an interpolation hybrid: two slots of one iteration each)
    int last = ipol ? 1 : nf - 1;
    while (!ipol && last > 0 && sc->formulas[last].iterations == 0) --last;
    p.n_slots = last + 1;
    FormulaInfo fi[MB3D_HIP_MAX_SLOTS];
    for (int i = 0; i < p.n_slots; i++) {
        const mb3d_hip_formula &f = sc->formulas[i];
        fi[i] = info_of(f);
        p.slot_kind[i] = f.kind;
        p.slot_iters[i] = ipol ? 1 : (f.iterations < 0 ? -f.iterations : f.iterations);
        p.slot_uncounted[i] = (!ipol && f.iterations < 0) || fi[i].de_option < 0 || fi[i].de_option > 20;
        p.slot_power[i] = f.power;
        p.slot_z_mul[i] = f.z_mul;
        p.slot_scale[i] = f.scale;
        p.slot_min_r[i] = f.min_r;
        p.slot_fold[i] = f.fold;
        if (const int pw = slot_options(f, p.slot_opt[i])) p.slot_power[i] = pw;
    }
    // a DE combination splits the hybrid in two parts, when both have slots
    // with iterations (src/calc.rs:421-431)
    const int dck = ipol ? 0 : sc->decomb;
    const bool dc = dck != 0 && sc->decomb_end1 < last && sc->decomb_start2 <= last;
    if (dc && (dck < 1 || dck > 6)) return set_error(MB3D_HIP_ERR_INVALID, "unknown DE combination");
    const int end1 = dc ? std::min(sc->decomb_end1, last) : last;
    p.end_to = end1;
    int rf = std::min(sc->repeat_from, end1);
    while (rf > 0 && p.slot_iters[rf] <= 0) --rf;
    p.repeat_from = rf;

    auto combined = [&](int from, int to) {
        int de = -1;
        for (int i = from; i <= to; i++)
            if (p.slot_iters[i] != 0 && fi[i].de_option >= 0 && fi[i].de_option != de) de = combine_de(de, fi[i].de_option);
        return de;
    };
    int de_option = combined(0, end1);
    // the second part: its slots, DE option and repeat slot (src/calc.rs:487-520)
    int de2 = -1;
    if (dc) {
        p.decomb = dck;
        p.start2 = std::min(sc->decomb_start2, last);
        p.end2 = std::min(std::max(sc->decomb_end2, p.start2), last);
        de2 = combined(p.start2, p.end2);
        if (de2 < 0)
            return set_error(MB3D_HIP_ERR_INVALID, "DE combination: the second formula part has no valid DE option");
        if (dck == 6) return set_error(MB3D_HIP_ERR_INVALID, "DE combination 'mix': the second hybrid part must be dIFS");
        int r2 = std::min(std::max(sc->decomb_repeat2, p.start2), p.end2);
        while (r2 > p.start2 && p.slot_iters[r2] <= 0) --r2;
        p.repeat2 = r2;
        p.max_it2 = sc->decomb_iterations2;
        p.de_option2 = de2;
        p.is_custom_de2 = de2 == 2 || de2 == 5 || de2 == 6 || de2 == 11 || de2 == 20;
        p.four_d2 = de2 >= 4 && de2 <= 6;
    }
    if (ipol) {
        // ipol_de_option (src/calc.rs:255-307): only the formulas whose raw
        // weight is positive (as the sign of its f32 bits) count
        de_option = -1;
        for (int x = 0; x < 2; x++) {
            int32_t bits;
            std::memcpy(&bits, &sc->interpolation_w[x], sizeof bits);
            const int de = fi[x].de_option;
            if (bits <= 0 || de < 0 || de > 20 || de == de_option) continue;
            const bool difs_mix = (de_option >= 0 && de_option < 20 && de > 19) || (de_option > 19 && de < 20);
            if (difs_mix) return set_error(MB3D_HIP_ERR_INVALID, "formula option is not valid");
            de_option = combine_de(de_option, de);
        }
        if (de_option > 19) return set_error(MB3D_HIP_ERR_INVALID, "dIFS formulas do not work in an interpolation hybrid");
        if (de_option < 0) return set_error(MB3D_HIP_ERR_INVALID, "formula option is not valid");
    }
    if (sc->de_mode == MB3D_HIP_DE_ANALYTIC_P8) {
        if (ipol || p.n_slots != 1 || sc->formulas[0].kind != MB3D_HIP_FORMULA_INTPOW)
            return set_error(MB3D_HIP_ERR_INVALID, "analytic p8 DE needs one Integer Power slot");
        de_option = 32;  // extension: running derivative
    } else if (sc->de_mode != MB3D_HIP_DE_REFERENCE) {
        return set_error(MB3D_HIP_ERR_INVALID, "unknown DE mode");
    }
    const bool custom = de_option == 2 || de_option == 5 || de_option == 6 || de_option == 11 || de_option == 20 ||
                        de_option == 32;
    p.de_option = de_option;
    p.is_custom_de = custom;
    for (int i = 0; i < p.n_slots; i++) p.slot_ade[i] = dc && i > end1 ? p.is_custom_de2 : custom;
    p.four_d = de_option >= 4 && de_option <= 6;  // HybridMode::Alt4D / Ipol4D (src/calc.rs:475-484)
    p.ipol = ipol;
    p.ipol_w[0] = ipw[0];
    p.ipol_w[1] = ipw[1];
    p.do_julia = sc->julia != 0;
    std::copy(sc->julia_c, sc->julia_c + 4, p.julia_c);
    smatrix4(sc->rot_4d, p.smatrix4);

    // DE scale, colour offset and smooth-iteration logs (src/calc.rs:534-600)
    const int max_it = sc->iterations;
    for (float &h : p.fhln) h = (float)(1.0 / std::log(2.0));
    // the sums of each part: the second part's slots are those after the
    // first part's last
    double sums[2][4] = {{0, 0, 0, 0}, {0, 0, 0, 0}};
    for (int i = 0; i < p.n_slots; i++) {
        if (p.slot_iters[i] == 0) continue;
        const bool second = dc && i > end1;
        double &cnt_all = sums[second][0], &scale_sum = sums[second][1], &abs_first = sums[second][2],
               &cnt_ade = sums[second][3];
        const bool custom_i = second ? p.is_custom_de2 != 0 : custom;
        // an interpolation hybrid weighs each formula by its weight and counts it once
        const double its = ipol ? (double)ipw[i] : (double)p.slot_iters[i];
        const double cnt = ipol ? 1.0 : its;
        if (ipol || !p.slot_uncounted[i]) {
            const double z1 = fi[i].si_pow > 1.0 ? fi[i].si_pow : fi[i].first_option;
            p.fhln[i] = (float)(1.0 / std::log(std::fmax(2.0, std::fabs(z1))));
        }
        const int j = fi[i].de_option;
        if (custom_i && (ipol || !p.slot_uncounted[i]) && (j == 5 || j == 11)) {
            cnt_all += cnt;
            cnt_ade += cnt;
            double x2 = fi[i].first_option;
            if (x2 == 0.0) x2 = 1e-40;
            scale_sum += x2 < 0.0 ? 0.65 * its : (x2 * 1.2 + 1.0) / x2 * its;
            abs_first += std::fabs(x2) * its;
        } else if (j >= 0 || ipol) {
            cnt_all += cnt;
            scale_sum += (custom_i ? 1.0 : fi[i].de_scale) * its;
        }
    }
    double de_scale0 = 1.0, de_scale2 = 1.0;
    if (sums[0][0] > 0.0 && de_option != 20) de_scale0 = ipol ? sums[0][1] : sums[0][1] / sums[0][0];
    if (dc && sums[1][0] > 0.0 && de2 != 20) de_scale2 = sums[1][1] / sums[1][0];
    // (with a DE combination the colour offset is the second part's)
    const double *cs = sums[dc ? 1 : 0];
    float col_plus = 0.0f;
    if (cs[0] > 0.0 && cs[3] > 0.0) col_plus = (float)(std::pow(std::fabs(cs[2] / cs[3]), 0.25) * 40.0 - 49.0);

    // mctColVarDEstopMul (src/calc.rs:594-650): per part; a DE combination
    // takes the part with more counted iterations
    float col_var = 0.6f;
    {
        double nums[2] = {0, 0}, dens[2] = {0, 0};
        int uncounted[2] = {0, 0};
        bool part2_counted = false;
        for (int i = 0; i < p.n_slots; i++) {
            const int part = dc && i > end1 ? 1 : 0;
            if (!ipol && !(p.slot_iters[i] > 0 && !p.slot_uncounted[i])) {
                if (p.slot_iters[i] > 0) uncounted[part]++;
                continue;
            }
            part2_counted |= part == 1;
            double &num = nums[part], &den = dens[part];
            const double its = ipol ? (double)ipw[i] : (double)p.slot_iters[i];
            const int de = fi[i].de_option;
            if (de == 0 || de == 4) {
                num += its * 0.6;
                den += its;
            } else if (((de >= 1 && de <= 3) || (de >= 5 && de <= 40)) && fi[i].first_option_double) {
                double v = fi[i].first_option;
                if (v > 1.0) {
                    if (v < 1.2) v = 1.1 + (v - 1.0) * (v - 1.0) * 2.5;
                    num += its * 0.5 / (std::log(v) + 0.03);
                    den += its;
                } else if (v < -1.0) {
                    if (v > -1.2) v = -1.1 - (v + 1.0) * (v + 1.0) * 2.5;
                    num += its * 0.5 / (std::log(-v) + 0.03);
                    den += its;
                }
            }
        }
        if (!part2_counted) {
            if (dens[0] > 1e-3) col_var = (float)(nums[0] / dens[0]);
        } else {
            col_var = dens[0] > 0.0 ? (float)(nums[0] / dens[0]) : 0.6f;
            const double z1 = dens[1], z2 = dens[0];
            if ((z1 >= z2 && z1 > 1e-3) || (uncounted[1] > uncounted[0] && std::fabs(z1 - z2) < 1e-3 && z1 > 1e-3))
                col_var = (float)(nums[1] / dens[1]);
            else if (z2 < 1e-3)
                col_var = 0.6f;
        }
    }

    // bailout (Scene::effective_rstop, src/scene.rs:799-809)
    double rstop = sc->rstop;
    if (!(rstop > 0.0)) {
        rstop = 0.0;
        for (int i = 0; i < nf; i++)
            if (sc->formulas[i].iterations > 0) rstop = std::fmax(rstop, info_of(sc->formulas[i]).default_rstop);
        rstop = std::fmax(rstop, 1.0001);
    }
    const double sw = 2.1345 / (sc->zoom * W);  // Scene::step_width
    const double fov = sc->fov_y * kRadPerDeg;
    const float zsd0 = (float)std::fmax(sc->z_step_div, 0.0001);
    const float de_stop = (float)std::fmax(sc->de_stop, 0.001);
    const double half = std::fmin(std::fmax(fov * 0.5, 0.01), 1.5);
    p.pl_optic_z = (float)(std::cos(half) * half / std::sin(half));
    p.fovx_off = 0.5f * (float)sc->width;
    p.fovx_mul = (float)(fov / Hh);

    // CalcPPZvals: the depth code's constants
    const double zcorr = std::sin(std::fmax(sc->fov_y * kRadPerDeg, 1.0) / Hh);
    const double depth_range = sc->z_end - sc->z_start;
    const double zc_mul = 32767.0 * 256.0 / (std::sqrt(depth_range * zcorr / sw + 1.0) - 0.999999999);
    const double zend = std::fmax(1e-10, depth_range / sw);
    const float mh04zsd = (float)((double)std::max(sc->width, sc->height) * 0.5 * std::sqrt((double)zsd0 + 0.001) *
                                  std::fmax(0.01, sc->raystep_limiter));
    // GetDEstopFactor
    float dsf = 0.0f;
    if (sc->vary_de_stop_on_fov) {
        const double ze = std::fmax(1e-16, depth_range / sw);
        const double a = 0.01 * Hh / (std::sin(0.01) * std::fmax(1.0 / 65535.0, sc->fov_y * kRadPerDeg));
        const double b = sw * (a + ze) / a;
        dsf = (float)((b - sw) / (sw * std::fmax(1e-6, ze)));
    }
    // camera: rows scaled to the step width, start of pixel (0, 0)
    const double z1 = (sc->z_start - sc->mid[2]) / sw;
    for (int i = 0; i < 3; i++) {
        const double *r = sc->vgrads[i];
        const double d = sw / std::sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
        for (int j = 0; j < 3; j++) p.vgrads[i][j] = r[j] * d;
    }
    for (int k = 0; k < 3; k++)
        p.ystart[k] = sc->mid[k] + z1 * p.vgrads[2][k] - (Hh * 0.5) * p.vgrads[1][k] - (W * 0.5) * p.vgrads[0][k];

    float de_offset = std::fmin(de_stop * 0.1f, 0.004f);
    p.d_de_scale = custom ? (float)((double)(float)de_scale0 / sw) : (float)de_scale0 * de_offset;
    if (dc) {
        p.d_de_scale2 = p.is_custom_de2 ? (float)((double)(float)de_scale2 / sw) : (float)de_scale2 * de_offset;
        p.decomb_smooth = (double)std::fmax(sc->decomb_smooth, 1e-30f) / sw;
        // the second part placed as an object of its own (Place, src/calc.rs:748-752)
        const double *o = sc->part2_offset, *r = sc->part2_rotation;
        const bool moved = o[0] != 0.0 || o[1] != 0.0 || o[2] != 0.0 || r[0] != 0.0 || r[1] != 0.0 || r[2] != 0.0 ||
                           sc->part2_scale != 1.0;
        if (sc->part2_placed && moved) {
            double a[3], sn[3], cs[3];
            for (int k = 0; k < 3; k++) {
                a[k] = r[k] * (kPi / 180.0);
                sn[k] = std::sin(a[k]);
                cs[k] = std::cos(a[k]);
            }
            // build_rot_matrix (src/math.rs:466-475), rows scaled to 1 / scale
            const double m[3][3] = {{cs[1] * cs[2], -cs[1] * sn[2], sn[1]},
                                    {sn[0] * sn[1] * cs[2] + cs[0] * sn[2], cs[0] * cs[2] - sn[0] * sn[1] * sn[2],
                                     -sn[0] * cs[1]},
                                    {sn[0] * sn[2] - cs[0] * sn[1] * cs[2], cs[0] * sn[1] * sn[2] + sn[0] * cs[2],
                                     cs[0] * cs[1]}};
            const double len = 1.0 / std::fmax(sc->part2_scale, 1e-300);
            for (int i = 0; i < 3; i++) {
                const double d = len / std::sqrt(m[i][0] * m[i][0] + m[i][1] * m[i][1] + m[i][2] * m[i][2]);
                for (int k = 0; k < 3; k++) p.place_m[i][k] = m[i][k] * d;
                p.place_offset[i] = o[i];
            }
            p.place_scale = sc->part2_scale;
            p.placed = 1;
        }
    }
    p.de_offset006 = de_offset * 0.06f;
    p.de_offset = (float)((double)de_offset * sw);

    p.z_step_div = zsd0;
    p.ms_de_sub = 0.0f;
    if (sc->step_sub_de_stop) {
        float s = zsd0;
        s = s * s + (1.2f * s) * (1.0f - s);
        p.z_step_div = s;
        p.ms_de_sub = std::fmin(std::sqrt(s), 0.9f);
    }
    p.ln_rstop = (float)std::log(std::log(rstop));
    p.d_rstop = (double)(float)(rstop * rstop);
    p.rstop3d = (double)(float)((p.d_rstop * p.d_rstop) * 64.0);
    // the iterations of both parts of a DE combination (iMaxitF2)
    const int its_all = !dc ? max_it : (dck > 5 ? max_it + sc->decomb_iterations2 : std::max(max_it, sc->decomb_iterations2));
    p.mcts_m = 32767.0f / (float)std::max(its_all + 1, 1);
    p.d_col_plus = col_plus + (float)its_all * 0.1f;
    p.col_var_de_stop_mul = col_var;

    p.sm_normals = std::min(std::max(sc->smooth_normals, 0), 8);
    p.de_stop = de_stop;
    p.de_add_steps = sc->bin_search_steps;
    p.fov_y = fov;
    p.max_it = max_it;
    p.min_it = std::min(sc->min_iterations, max_it);
    p.color_option = std::min(sc->color_option, 5);
    p.color_on_it = std::min(std::max(sc->color_on_it, 0), 255);  // (a u8 in CalcParams, src/calc.rs:27)
    p.mct_color_mul = (float)(sc->color_mul * 512.0);
    p.dfog_on_it = sc->dfog_on_it;
    p.zcorr = zcorr;
    p.zc_mul = zc_mul;
    p.zend = zend;
    p.step_width = sw;
    p.first_step_random = sc->first_step_random != 0;
    p.mh04zsd = mh04zsd;
    p.de_stop_factor = dsf;
    p.normals_on_de = sc->normals_on_de || custom;
    p.step_cap = std::max(sc->step_cap, 0);
    *out = p;
    return MB3D_HIP_OK;
}

extern "C" int32_t mb3d_hip_supported(const mb3d_hip_params *p, char *reason, size_t reason_len) {
    std::string why = mb3d::unsupported_reason(p);
    if (why.empty()) return MB3D_HIP_OK;
    if (reason && reason_len) std::snprintf(reason, reason_len, "%s", why.c_str());
    mb3d::set_error(MB3D_HIP_UNSUPPORTED, why.c_str());
    return MB3D_HIP_UNSUPPORTED;
}

extern "C" int32_t mb3d_hip_pixel_seed(int64_t y, int64_t x) {
    // the row seed advanced x steps of s -> 214013 s + 2531011, by squaring
    uint32_t s = (uint32_t)(int32_t)(0x24563487LL + (y + 1) * 0x324594A1LL);
    uint32_t mul = 214013u, add = 2531011u, acc_mul = 1u, acc_add = 0u;
    for (uint64_t n = x > 0 ? (uint64_t)x : 0; n; n >>= 1) {
        if (n & 1) {
            acc_add = acc_add * mul + add;
            acc_mul *= mul;
        }
        add = add * mul + add;
        mul *= mul;
    }
    return (int32_t)(acc_mul * s + acc_add);
}
